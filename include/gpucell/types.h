/* gpucell/types.h -- plain-old-data layouts shared by the host, the HIP
 * kernels and the C-ABI of libgpucell_hip.so.
 *
 * The structs keep the field names, order, sizes and alignment of the
 * reference's TypesDef (src/TypesDef.cuh:6-78), so arrays allocated by
 * GPUCell-style host code can be handed to this library unchanged.
 * Sizes/alignments are pinned by the static asserts at the end.
 *
 * Everything past Structconst (gc_ext_consts) has no counterpart in the
 * reference: it carries the choices the reference hard-wired (periodic axes,
 * forcing pattern, membrane-law constants, thermal-noise counter).
 */
#ifndef GPUCELL_TYPES_H
#define GPUCELL_TYPES_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* src/TypesDef.cuh:6-10 */
typedef struct { int x, y; } tint2;
/* src/TypesDef.cuh:12-16 */
typedef struct { int x, y, z; } tint3;
/* src/TypesDef.cuh:18-22; membrane edge record (Particle.cuh:35): x,y shared
 * edge vertices, z,k the opposite vertex of triangle v resp. w, v,w the two
 * triangles.  All ids are 1-based (see the loader, Particle.cu:268-284). */
typedef struct { int x, y, z, k, v, w; } tint6;
/* src/TypesDef.cuh:24-28 (the reference declares the same record twice). */
typedef struct { int x, y, z, k, v, w; } int6;
/* src/TypesDef.cuh:31-35 (unused by the reference's step). */
typedef struct { float x, y, z, w, v; } tfloat5;
/* src/TypesDef.cuh:37-41; Velrhop = (vx, vy, vz, rho). */
typedef struct { float x, y, z, w; } tfloat4;
/* src/TypesDef.cuh:43-47 */
typedef struct { float x, y, z; } tfloat3;

/* src/TypesDef.cuh:49-78 -- the 96-byte parameter block that the reference
 * uploads once with cupar::ConstsDataUp (Particle.cu:654-664).  Id ranges
 * (Particle.cu:432-435):
 *   [0, nBound2Start)            Bound1 walls (repulsion only)
 *   [nBound2Start, nFluidStart)  Bound2 walls (SDPD, no-slip velocity)
 *   [nFluidStart, nCellStart)    fluid
 *   [nCellStart, nTotalPar)      membrane vertices (nCellPar of them)      */
typedef struct {
    unsigned nFluidStart;
    unsigned nBound2Start;
    unsigned nCellStart;
    unsigned nTotalPar;
    unsigned nCellPar;
    long int iSeed;        /* thermal-noise seed */
    float CutRadius;       /* SDPD kernel support h (= cube edge lower bound) */
    float Mass;
    float SoundSpeed1;     /* fluid and wall sound speed */
    float SoundSpeed2;     /* membrane-vertex sound speed */
    float DenInit;         /* rest mass density rho0 */
    float VVsigma;         /* Groot-Warren lambda */
    float ShearVis;        /* eta */
    float BulkVis;         /* zeta */
    float ConstExtForce;   /* body force per particle (already times Mass) */
    float TimeStep;
    float surEnergy;       /* Morse D_e */
    float MaxCut;          /* Morse cut-off */
    float zeroLen;         /* Morse r0 */
    float scaFac;          /* Morse beta */
    float Temp;            /* k_B T */
} Structconst;

/* Forcing patterns for SetFluidExtForce. */
enum {
    GC_FORCE_UNIFORM = 0,          /* +x on every moving particle */
    GC_FORCE_REVERSE_POISEUILLE = 1 /* +x where y < force_split, -x above */
};

/* Settings the reference fixes in code; set with gc_SetExtConsts.
 * Defaults (gc_DefaultExtConsts) reproduce the reference's domain convention:
 * x periodic, y and z bounded by wall particles (Particle.cu:551-559). */
typedef struct {
    int periodic[3];        /* 1 = periodic axis (no cube padding) */
    int force_mode;         /* GC_FORCE_* */
    float force_split;      /* y plane of the reverse-Poiseuille split */
    int noise_on;           /* 1 = SDPD random term on */
    float wlc_ratio;        /* l_max / l0 of the WLC springs (x0 = 1/ratio) */
    float pow_index;        /* POW exponent m */
    float rep_scale;        /* Bound1 repulsion: multiple of c1^2 rho0 */
    float wall_split;       /* moving walls (plane Couette shear): Bound2 walls with */
    float wall_u_lo;        /* y < wall_split move along +x at wall_u_lo, the others */
    float wall_u_hi;        /* at wall_u_hi (positions stay fixed; 0 = at rest)     */
    uint64_t noise_counter; /* counter of the next SDPD force evaluation */
} gc_ext_consts;

#ifdef __cplusplus
} /* extern "C" */
#endif

#if defined(__cplusplus)
#define GC_STATIC_ASSERT(c, m) static_assert(c, m)
#define GC_ALIGNOF(t) alignof(t)
#else
#define GC_STATIC_ASSERT(c, m) _Static_assert(c, m)
#define GC_ALIGNOF(t) _Alignof(t)
#endif

GC_STATIC_ASSERT(sizeof(tint2) == 8 && GC_ALIGNOF(tint2) == 4, "tint2 layout");
GC_STATIC_ASSERT(sizeof(tint3) == 12 && GC_ALIGNOF(tint3) == 4, "tint3 layout");
GC_STATIC_ASSERT(sizeof(int6) == 24 && GC_ALIGNOF(int6) == 4, "int6 layout");
GC_STATIC_ASSERT(sizeof(tint6) == 24, "tint6 layout");
GC_STATIC_ASSERT(sizeof(tfloat3) == 12 && GC_ALIGNOF(tfloat3) == 4, "tfloat3 layout");
GC_STATIC_ASSERT(sizeof(tfloat4) == 16 && GC_ALIGNOF(tfloat4) == 4, "tfloat4 layout");
GC_STATIC_ASSERT(sizeof(tfloat5) == 20, "tfloat5 layout");
GC_STATIC_ASSERT(sizeof(Structconst) == 96 && GC_ALIGNOF(Structconst) == 8, "Structconst layout");
GC_STATIC_ASSERT(offsetof(Structconst, nCellPar) == 16, "Structconst.nCellPar");
GC_STATIC_ASSERT(offsetof(Structconst, iSeed) == 24, "Structconst.iSeed");
GC_STATIC_ASSERT(offsetof(Structconst, CutRadius) == 32, "Structconst.CutRadius");
GC_STATIC_ASSERT(offsetof(Structconst, Temp) == 88, "Structconst.Temp");
GC_STATIC_ASSERT(sizeof(gc_ext_consts) == 56, "gc_ext_consts layout");
GC_STATIC_ASSERT(offsetof(gc_ext_consts, noise_counter) == 48, "gc_ext_consts.noise_counter");

#endif /* GPUCELL_TYPES_H */
