"""Material parameters and conserved-variable construction (host side).

`Par` is the ctypes mirror of `gpr_par` (include/gpr_par.h), itself a plain-C
restatement of the reference's `struct Par` (code/cpp/src/system/objects.h:22-98).
Attribute access works with the reference's Unicode field names (``MP.ρ0``,
``MP.μ``, ``MP.cα2`` ...) as well as the ASCII ones, so setup code written for
``GPRpy.classes.Par`` (code/cpp/bindings/bindings.cpp:115-171) runs unchanged.

`material_params` restates code/python/gpr/misc/objects.py:184-267 (and the
field assignments of `params`, objects.py:44-181) and `Cvec` restates
code/python/gpr/misc/structures.py:104-126 with the Mie-Gruneisen energy of
code/python/gpr/vars/eos.py:41-57 and mg.py:201-208.
"""

import ctypes
import math

import numpy as np

EOS_CODES = {"vac": -1, "sg": 0, "smg": 1, "jwl": 2, "cc": 3, "gr": 4}

_DOUBLE_FIELDS = [
    "rho0", "Tref", "cv", "pINF", "Gamma0", "c02", "s", "alpha", "beta",
    "gamma", "A", "B", "R1", "R2", "b02", "mu", "tau0", "sigmaY", "n",
    "calpha2", "kappa", "T0",
]

_UNICODE = {
    "ρ0": "rho0", "Γ0": "Gamma0", "α": "alpha", "β": "beta", "γ": "gamma",
    "μ": "mu", "τ0": "tau0", "σY": "sigmaY", "cα2": "calpha2", "κ": "kappa",
    "δp": "dp",
}


class Par(ctypes.Structure):
    """ctypes layout of `gpr_par` (include/gpr_par.h)."""

    _fields_ = ([(f, ctypes.c_double) for f in _DOUBLE_FIELDS]
                + [("dp", ctypes.c_double * 3)]
                + [(f, ctypes.c_int32) for f in
                   ("EOS", "SOLID", "POWER_LAW", "BINGHAM", "REACTION", "pad_")])

    def __getattr__(self, name):
        alias = _UNICODE.get(name)
        if alias is None:
            raise AttributeError(name)
        val = getattr(self, alias)
        if alias == "dp":
            return np.array(val[:], dtype=np.float64)
        return val

    def __setattr__(self, name, value):
        name = _UNICODE.get(name, name)
        if name == "dp":
            value = (ctypes.c_double * 3)(*[float(v) for v in np.asarray(value).ravel()])
        elif name in ("SOLID", "POWER_LAW", "BINGHAM", "MULTI"):
            if name == "MULTI":
                if value:
                    raise ValueError("MULTI mixtures are not supported on this path")
                return
            value = int(bool(value))
        super().__setattr__(name, value)

    def copy(self):
        c = Par()
        ctypes.memmove(ctypes.byref(c), ctypes.byref(self), ctypes.sizeof(Par))
        return c

    @classmethod
    def from_dict(cls, d):
        """The inverse of as_dict."""
        MP = cls()
        for k, v in d.items():
            setattr(MP, k, v)
        return MP

    def as_dict(self):
        d = {f: getattr(self, f) for f in _DOUBLE_FIELDS}
        d["dp"] = list(self.dp)
        for f in ("EOS", "SOLID", "POWER_LAW", "BINGHAM", "REACTION"):
            d[f] = getattr(self, f)
        return d


# Params is the base struct in the reference; on this path the two coincide.
Params = Par


def material_params(EOS, ρ0, cv=None, Tref=None, α=None, β=None, γ=None,
                    pINF=None, c0=None, Γ0=None, s=None, A=None, B=None,
                    R1=None, R2=None, b0=None, μ=None, τ0=None, σY=None, n=None,
                    cα=None, κ=None, Pr=None, δp=None):
    """Restates objects.py:184-267 for the single-material, non-reactive path."""
    MP = Par()
    MP.REACTION = -1
    if EOS == "vac":
        MP.EOS = -1
        return MP
    if Tref is None:
        Tref = 0
    if γ and pINF is None:
        pINF = 0
    if n:
        POWER_LAW = True
        SOLID = bool(σY)
        BINGHAM = False
    else:
        n = 1
        POWER_LAW = False
        SOLID = not μ
        BINGHAM = bool(μ and σY)
    if β is None:
        β = 0
    if Pr:
        κ = μ * γ * cv / Pr

    # params(), objects.py:44-181
    MP.EOS = EOS_CODES[EOS]
    MP.rho0 = ρ0
    if cv:
        MP.cv = cv
        MP.Tref = Tref
    if EOS == "sg":
        MP.gamma = γ
        MP.pINF = pINF
    elif EOS == "smg":
        MP.Gamma0 = Γ0
        MP.c02 = c0 ** 2
        MP.s = s
    elif EOS == "gr":
        MP.c02 = c0 ** 2
        MP.alpha = α
        MP.beta = β
        MP.gamma = γ
    elif EOS in ("jwl", "cc"):
        MP.Gamma0 = Γ0
        MP.A = A
        MP.B = B
        MP.R1 = R1
        MP.R2 = R2
    if b0:
        MP.b02 = b0 ** 2
        MP.beta = β
        MP.POWER_LAW = POWER_LAW
        MP.SOLID = SOLID
        MP.BINGHAM = BINGHAM
        if μ:
            MP.mu = μ
        if n:
            MP.n = n
        if σY:
            MP.sigmaY = σY
        if τ0:
            MP.tau0 = τ0
    if cα:
        MP.calpha2 = cα ** 2
        MP.kappa = κ
    MP.dp = np.zeros(3) if δp is None else δp
    return MP


# --- Mie-Gruneisen reference curves (code/python/gpr/vars/mg.py) -------------

def _Γ_MG(ρ, MP):
    if MP.EOS == 0:
        return MP.gamma - 1
    if MP.EOS == 1:
        return MP.Gamma0 * MP.rho0 / ρ
    if MP.EOS == 4:
        return MP.gamma
    if MP.EOS in (2, 3):
        return MP.Gamma0
    raise ValueError("EOS not recognized")


def _p_ref(ρ, MP):
    ρ0 = MP.rho0
    if MP.EOS == 0:
        return -MP.pINF
    if MP.EOS == 1:
        if ρ > ρ0:
            tmp = 1 / ρ0 - MP.s * (1 / ρ0 - 1 / ρ)
            return MP.c02 * (1 / ρ0 - 1 / ρ) / tmp ** 2
        return MP.c02 * (ρ - ρ0)
    if MP.EOS == 4:
        tmp = (ρ / ρ0) ** MP.alpha
        return MP.c02 * ρ / MP.alpha * (tmp - 1) * tmp
    v_ = ρ0 / ρ
    if MP.EOS == 2:
        return MP.A * math.exp(-MP.R1 * v_) + MP.B * math.exp(-MP.R2 * v_)
    if MP.EOS == 3:
        return MP.A * v_ ** (-MP.R1) - MP.B * v_ ** (-MP.R2)
    raise ValueError("EOS not recognized")


def _e_ref(ρ, MP):
    ρ0 = MP.rho0
    if MP.EOS == 0:
        return MP.pINF / ρ
    if MP.EOS == 1:
        if ρ > ρ0:
            return 0.5 * _p_ref(ρ, MP) * (1 / ρ0 - 1 / ρ)
        return 0.0
    if MP.EOS == 4:
        tmp = (ρ / ρ0) ** MP.alpha
        return MP.c02 / (2 * MP.alpha ** 2) * (tmp - 1) ** 2
    v_ = ρ0 / ρ
    if MP.EOS == 2:
        return (MP.A / (ρ0 * MP.R1) * math.exp(-MP.R1 * v_)
                + MP.B / (ρ0 * MP.R2) * math.exp(-MP.R2 * v_))
    if MP.EOS == 3:
        return (-MP.A / (ρ0 * (1 - MP.R1)) * (v_ ** (1 - MP.R1) - 1)
                + MP.B / (ρ0 * (1 - MP.R2)) * (v_ ** (1 - MP.R2) - 1))
    raise ValueError("EOS not recognized")


def internal_energy(ρ, p, MP):
    """mg.py:201-208: e = e_ref + (p - p_ref) / (ρ Γ)."""
    return _e_ref(ρ, MP) + (p - _p_ref(ρ, MP)) / (ρ * _Γ_MG(ρ, MP))


def _devGsq(A):
    G = A.T @ A
    devG = G - np.trace(G) / 3 * np.eye(3)
    return float(np.sum(devG * devG))


def total_energy(ρ, p, v, A, J, MP):
    """eos.py:41-57 with E_2A = C0/4 |dev G|^2 (C0 = (1 - β/Γ) cs^2, cs^2 = b0^2 (ρ/ρ0)^β)."""
    E = internal_energy(ρ, p, MP)
    if A is not None:
        cs2 = MP.b02 * (ρ / MP.rho0) ** MP.beta
        C0 = (1 - MP.beta / _Γ_MG(ρ, MP)) * cs2
        E += C0 / 4 * _devGsq(np.asarray(A, dtype=np.float64))
    if J is not None:
        E += MP.calpha2 * float(np.dot(J, J)) / 2
    E += float(np.dot(v, v)) / 2
    return E


def Cvec(ρ, p, v, MP, A=None, J=None):
    """structures.py:104-126: conserved vector [ρ, ρE, ρv, A (row-major), ρJ]."""
    NV = 5 + 9 * (A is not None) + 3 * (J is not None)
    Q = np.zeros(NV)
    v = np.asarray(v, dtype=np.float64)
    Q[0] = ρ
    Q[1] = ρ * total_energy(ρ, p, v, A, J, MP)
    Q[2:5] = ρ * v
    if A is not None:
        Q[5:14] = np.asarray(A, dtype=np.float64).ravel()
    if J is not None:
        Q[14:17] = ρ * np.asarray(J, dtype=np.float64)
    return Q
