/*
 * CPU oracle for the seismic C-PML time loop -- TEST INFRASTRUCTURE ONLY.
 * See cpml_oracle.h for the contract and the reference citations.
 * Built by oracle/Makefile into oracle/liboracle.so.
 */
#include "cpml_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ADE8:263-266 */
const double oracle_holberg8[4] = {1.231666, -1.041182e-1, 2.063707e-2, -3.570998e-3};

#define REAL double
#define SFX d
#define LIT(x) x
#define EXP exp
#define LOG log
#define SIN sin
#define COS cos
#define SQRT sqrt
#define FABS fabs
#include "cpml_oracle_body.h"
#undef REAL
#undef SFX
#undef LIT
#undef EXP
#undef LOG
#undef SIN
#undef COS
#undef SQRT
#undef FABS

#define REAL float
#define SFX f
#define LIT(x) x##f
#define EXP expf
#define LOG logf
#define SIN sinf
#define COS cosf
#define SQRT sqrtf
#define FABS fabsf
#include "cpml_oracle_body.h"
