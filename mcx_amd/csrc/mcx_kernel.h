/*
 * mcx_kernel.h -- launch parameters of the transport kernel.
 *
 * KParam is passed by value as the kernel argument block, so every field
 * is wave-uniform and lands in scalar registers; it replaces the
 * reference's __constant__ MCXParam (src/mcx_core.h:138-198) and the
 * gproperty table (src/mcx_core.cu:180).  The property/detector table is
 * read from global memory once per workgroup and kept in LDS.
 */
#ifndef MCX_AMD_KERNEL_H
#define MCX_AMD_KERNEL_H

#include <stdint.h>

/* tally precision modes */
enum {
    MCX_TALLY_F32_SHADOW = 0, /* FP32 atomics + MAX_ACCUM shadow half (reference semantics) */
    MCX_TALLY_F32_HILO = 1,   /* FP32 hi + FP32 lo (TwoSum error), merged in FP64 */
    MCX_TALLY_F64 = 2,        /* FP64 no-return atomics */
    MCX_TALLY_F64_XCD = 3     /* FP64, one private copy per XCD, workgroup-scope atomics */
};

#define MCX_NUM_XCD 8

/* Hi/lo tally (mode 1): the FP32 hi half is kept within about one chunk of
 * zero by moving whole chunks to an integer count per cell; a power of two,
 * so the moves and the count are exact */
#define MCX_HILO_CHUNK 1024.0f

struct KParam {
    /* grid */
    uint32_t nx, ny, nz, dimxy, nvox;
    float maxidx_x, maxidx_y, maxidx_z;
    /* source */
    float srcpos[4], srcdir[4], srcparam1[4], srcparam2[4];
    int32_t srctype;
    /* time gates */
    float twin0, twin1, Rtstep, oneoverc0, minaccumtime;
    uint32_t maxgate;
    /* physics */
    float minenergy;
    uint32_t doreflect, isspecular, voidtime, maxvoidstep, gscatter, outputtype, save2pt, is2d;
    uint32_t bcface;  /* face f (x-,y-,z-,x+,y+,z+) boundary code in bits 4f..4f+3 */
    uint32_t bcdet;   /* bit f: face f of the domain is a detector (bc chars 7-12) */
    /* detected photons */
    uint32_t savedet, savedetflag, maxdetphoton, reclen, partialdata, maxmedia, detnum, medianum;
    uint32_t ntable; /* medianum + detnum rows of float4 */
    /* work split (src/mcx_core.cu:3014-3022) */
    int32_t threadphoton, oddphotons;
    uint32_t nthread;
    /* buffers */
    const void* media;            /* uint8 (label | det<<7), uint32 (label | DET_MASK) or property words */
    const float* gprop;           /* ntable x 4 floats */
    float* field;                 /* tally (FP32 modes) */
    float* field2;                /* shadow (mode 0) or lo (mode 1) */
    int32_t* fieldk;              /* mode 1: chunks of MCX_HILO_CHUNK moved out of the hi half */
    double* field64;              /* modes 2, 3 (mode 3: MCX_NUM_XCD copies of fieldlen) */
    uint64_t fieldlen;            /* voxels x gates of one tally copy */
    float* genergy;               /* 2*nthread */
    const uint32_t* seeds;        /* 4*nthread */
    float* detdata;               /* maxdetphoton*reclen */
    uint32_t* detcount;           /* [0]=detected, [1]=error exits, [2]=lanes done */
    uint64_t* counters;           /* optional [steps, deposits, scatters, launches] or NULL */
    const float* srcpattern;      /* pattern / pattern3d source intensities, or NULL */
    uint32_t issaveref;           /* tally escaping weight (negative) in the void voxel a photon leaves into */
    /* seed saving and replay */
    uint32_t saveseed;            /* store each detected photon's launch RNG state */
    uint32_t* seeddata;           /* 4 words per detected-photon slot */
    int32_t replay;               /* photons relaunched from saved RNG states */
    int32_t replaydet;            /* 0: all detectors, -1: one tally block per detector, k: detector k */
    const uint32_t* replayseed;   /* 4 words per replayed photon (u64 s0, s1, little endian) */
    const float* replayweight;    /* detected weight of each replayed photon */
    const float* replaytof;       /* time of flight of each replayed photon */
    const int32_t* replaydetid;   /* detector id of each replayed photon */
    int32_t replayindex;          /* 0: photon k of thread idx is idx*threadphoton + min(idx, oddphotons) + k;
                                     1: idx*threadphoton + min(idx, oddphotons - 1) + k (the reference's) */
    float omega;                  /* RF replay: modulation angular frequency (rad/s) */
    int32_t rfshift;              /* RF replay: gates from the real plane to the imaginary plane; 0 otherwise */
    uint32_t mediaformat;         /* 1, 2, 4: labels; 96-104: per-voxel property words */
    /* Hot-cell box: the tally cells of gates [0, hot_g) over voxels
     * [hot_x0, hot_x0 + hot_bx) x [hot_y0, ..) x [hot_z0, ..) are summed in
     * FP64 in each workgroup's LDS (at byte offset hot_off of the dynamic
     * LDS) and added to the tally once per workgroup at its end.  These
     * cells, next to the source, take the most deposits per cell, and
     * global atomics on one address serialise.  hot_cells == 0: off. */
    uint32_t hot_cells, hot_off, hot_g;
    uint32_t hot_x0, hot_y0, hot_z0, hot_bx, hot_by, hot_bz;
    /* division of a voxel index by dimxy and by nx (DivMagic) */
    uint32_t hot_mxy, hot_shxy, hot_mx, hot_shx;
    /* Blocked FP64 tally (tblock = 1; single-simulation FP64 tally): the
     * cell of voxel (x, y, z) in gate plane g is element
     *   8 * (((g * bnz + z/2) * bny + y/2) * bnx + x/2) + (x&1) + 2 (y&1) + 4 (z&1),
     * so the 8 cells of a 2 x 2 x 2 voxel block share one 64-byte line.
     * Each lane sums the deposits into its current block in LDS (9 words
     * per lane at byte offset blk_off: 8 FP32 cells and the block index);
     * when a lane moves to another block the wave adds the blocks left by
     * its lanes to the tally together, 8 lanes to a block, one atomic
     * instruction per 8 blocks (lane-id queue of 64 words per wave at
     * bq_off).  The device's atomic unit takes one 64-byte line per
     * request, so a block's cells cost one request instead of one each. */
    uint32_t tblock, bnx, bny, bnz, blk_off, bq_off;
    /* Replicas: virtual thread idx tallies into replica idx & nmaps_mask,
     * whose FP64 tally starts mapstride elements x replica on (0: one run) */
    uint32_t nmaps_mask;
    uint64_t mapstride;
    /* refill = 1: nlanes lanes run the nvthreads virtual threads, lane l
     * first thread l, then each the next unstarted one (detcount[3]) */
    uint32_t refill, nlanes, nvthreads;
};


/* Unsigned 32-bit division by an invariant d >= 1 as a multiply-high, two
 * shifts and two adds: Granlund and Montgomery, "Division by Invariant
 * Integers using Multiplication", PLDI 1994, Figure 4.1.  With
 * l = ceil(log2 d): m = floor(2^32 (2^l - d) / d) + 1, sh1 = min(l, 1),
 * sh2 = max(l - 1, 0); then for every 32-bit n, with t = mulhi(m, n),
 * n / d = (t + ((n - t) >> sh1)) >> sh2.  The shifts are packed as
 * sh1 | sh2 << 8. */
static inline void div_magic(uint32_t d, uint32_t* m, uint32_t* sh) {
    uint32_t l = 0;
    while (l < 32 && (1ull << l) < d) l++;
    *m = (uint32_t)((((1ull << l) - d) << 32) / d + 1);
    *sh = (l < 1 ? l : 1u) | ((l > 0 ? l - 1 : 0u) << 8);
}

/* One instance of a batched launch (its launch constants are in KParam) */
struct BatchInst {
    float* out;          /* merged, normalised FP32 fluence (fieldlen) */
    double* energy;      /* out: [0] escaped, [1] launched (FP64 sums over lanes), [2] normalizer */
    float unitinmm;      /* normalisation inputs, as the single-simulation path uses them */
    float vvox;
    float tstep;
    int32_t outputtype;
    int32_t normalize;   /* 0: raw tally */
    int32_t pad;
    float* cw;           /* time-integrated map (nvox), filled by mcx_batch_cw */
};

#endif