/* Material parameters of one GPR material, as a plain C struct.
 *
 * Field-for-field restatement of the reference's `struct Params` / `struct Par`
 * (code/cpp/src/system/objects.h:22-98) restricted to what the ADER-WENO step
 * reads on the single-material path.  ASCII names replace the reference's
 * Unicode identifiers (rho0 = ρ0, Gamma0 = Γ0, mu = μ, sigmaY = σY,
 * calpha2 = cα2, kappa = κ, dp = δp); the Python mirror exposes the Unicode
 * names as well (classes.Par).
 *
 * The layout is fixed (all doubles, then int32 flags) so that ctypes, the
 * reference-build driver and the device code agree on it byte for byte.
 */
#ifndef GPR_PAR_H
#define GPR_PAR_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* EOS codes, code/cpp/src/system/energy/mg.h:6-11 */
#define GPR_EOS_VACUUM (-1)
#define GPR_EOS_STIFFENED_GAS 0
#define GPR_EOS_SHOCK_MG 1
#define GPR_EOS_JWL 2
#define GPR_EOS_COCHRAN_CHAN 3
#define GPR_EOS_GODUNOV_ROMENSKI 4

typedef struct gpr_par {
  /* Params (objects.h:22-58) */
  double rho0;    /* ρ0   reference density                    */
  double Tref;    /* Tref reference temperature                */
  double cv;      /* cv   specific heat at constant volume     */
  double pINF;    /* pINF stiffened-gas pressure               */
  double Gamma0;  /* Γ0   Grüneisen coefficient                */
  double c02;     /* c0²  bulk sound speed squared             */
  double s;       /* s    Hugoniot slope                       */
  double alpha;   /* α    Godunov-Romenski exponent            */
  double beta;    /* β    shear-modulus density exponent       */
  double gamma;   /* γ    adiabatic index                      */
  double A;       /* JWL / Cochran-Chan A                      */
  double B;       /* JWL / Cochran-Chan B                      */
  double R1;      /* JWL / Cochran-Chan R1                     */
  double R2;      /* JWL / Cochran-Chan R2                     */
  double b02;     /* b0² = c_s² shear wave speed squared       */
  double mu;      /* μ    viscosity                            */
  double tau0;    /* τ0   plastic relaxation time              */
  double sigmaY;  /* σY   yield stress                         */
  double n;       /* n    power-law exponent                   */
  double calpha2; /* cα²  heat wave speed squared              */
  double kappa;   /* κ    heat conductivity                    */
  /* Par (objects.h:60-98) */
  double T0;
  double dp[3];   /* δp   body force                           */
  int32_t EOS;
  int32_t SOLID;
  int32_t POWER_LAW;
  int32_t BINGHAM;
  int32_t REACTION; /* only -1 (no reaction) is on this path   */
  int32_t pad_;
} gpr_par;

#ifdef __cplusplus
}
#endif

#endif /* GPR_PAR_H */
