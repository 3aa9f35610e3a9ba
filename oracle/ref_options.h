/* Build-time configuration for compiling the reference solver into oracle/_ref.
 *
 * Force-included (g++ -include) ahead of every reference translation unit.  It
 * defines the include guard OPTIONS_H that code/cpp/src/options.h:1-2 tests, so
 * the reference's own options header is skipped and these values are used
 * instead.  Every constant keeps the reference's value (options.h:5-35) except:
 *   N        polynomial nodes per dimension   (options.h:15 ships 2; P2 needs 3)
 *   THERMAL  heat-conduction variables on/off (options.h:6 ships false)
 *   LSET     level-set count                  (options.h:8 ships 2; single-material
 *                                              runs need 0, see iterator.cpp:114-117)
 * which come from -DGPR_REF_N / -DGPR_REF_THERMAL on the compile line.
 * This header is part of the oracle build recipe, not of the product. */
#ifndef OPTIONS_H
#define OPTIONS_H

#ifndef GPR_REF_N
#define GPR_REF_N 3
#endif
#ifndef GPR_REF_THERMAL
#define GPR_REF_THERMAL 0
#endif

const bool VISCOUS = true;
const bool THERMAL = (GPR_REF_THERMAL != 0);
const bool MULTI = false;
const int LSET = 0;

const bool NO_CORNERS = true;
const bool DESTRESS = false;

const int N = GPR_REF_N;
const bool ROTATE_DISTORTION = false;
const int RK_STEPS = 4;

const bool RIEMANN_STICK = false;
const bool RIEMANN_RELAXATION = true;
const double STAR_TOL = 1e-8;

const double LAMS = 1.;
const double LAMC = 1e5;
const double EPS = 1e-14;
const bool PRIM_RECONSTRUCT = false;

const int DG_IT = 50;
const double DG_TOL = 6e-6;

const int V = 5 + int(VISCOUS) * 9 + int(THERMAL) * 3 + int(MULTI) + LSET;
const int mV = 5 + int(VISCOUS) * 9 + int(THERMAL) * 3;

#endif
