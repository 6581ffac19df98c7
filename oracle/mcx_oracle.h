/*
 * mcx_oracle.h -- TEST INFRASTRUCTURE ONLY.
 *
 * A CPU (OpenMP) restatement of the reference's photon-transport loop,
 * used by tests/ as the checker and by bench.py's cpu_baseline leg as the
 * host-core baseline.  Nothing in the product (mcx_amd/) links or calls it.
 *
 * Inputs are the *preprocessed* quantities the reference's kernel sees
 * (src/mcx_core.cu:2853-2876 MCXParam, :3316-3321 gproperty): grid-unit
 * media properties, 0-based source/detector positions, column-major
 * uint32 labels with DET_MASK in bit 31, and the per-thread seed words.
 */
#ifndef MCX_ORACLE_H
#define MCX_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct OracleParams {
    /* domain */
    unsigned int nx, ny, nz;
    const unsigned int* media;   /* nx*ny*nz labels or words, x fastest, DET_MASK in bit 31 */
    const float* prop;           /* medianum x {mua,mus,g,n}, grid units */
    unsigned int medianum;
    const float* detpos;         /* detnum x {x,y,z,r}, 0-based grid units */
    unsigned int detnum;
    /* source */
    int srctype;
    float srcpos[4], srcdir[4], srcparam1[4], srcparam2[4];
    /* time gates */
    float twin0, twin1, Rtstep, oneoverc0, minaccumtime;
    unsigned int maxgate;
    /* physics flags */
    float minenergy;
    int isreflect;               /* template isreflect (cfg.isreflect or any r/m bc) */
    int doreflect;               /* cfg.isreflect */
    int isspecular, voidtime, maxvoidstep;
    unsigned int gscatter;
    int outputtype;              /* otFlux / otFluence / otEnergy / otL; replay: otJacobian, otWP, otDCS, otRF, otRFmus, otWLTOF, otWPTOF */
    int save2pt;
    unsigned char bc[12];        /* numeric bc codes and det flags */
    int is2d;
    /* detected photons */
    int savedet;                 /* cfg.issavedet (0,1,3) */
    unsigned int savedetflag, maxdetphoton;
    /* launch layout */
    unsigned int nthread;        /* virtual threads (RNG streams) */
    int threadphoton, oddphotons;
    const unsigned int* seeds;   /* 4*nthread seed words */
    unsigned int thread_begin, thread_end;  /* run streams [begin,end) */
    int ompthreads;              /* 0 = OpenMP default */
    /* seed saving and replay (src/mcx_core.cu:1205-1215, :2019-2047, :2208-2220) */
    int saveseed;                /* record each detected photon's launch RNG state */
    const unsigned int* replayseed;  /* NULL, or 4 words per photon: u64 s0, s1 little endian */
    const float* replayweight;
    const float* replaytof;
    const int* replaydetid;
    int replaydet;
    const float* srcpattern;     /* pattern / pattern3d source intensities (one pattern), or NULL */
    int issaveref;               /* diffuse reflectance into void voxels (src/mcx_core.cu:1124-1145) */
    int mediaformat;             /* 0-4: labels; 96-104: property words (96, 97: 2 words per voxel) */
    /* RF replay (otRF, otRFmus; src/mcx_core.cu:1700-1703, :2033-2062,
     * :2208-2213, :2254-2257): modulation angular frequency in rad/s; the
     * tally then holds a real plane and an imaginary plane of
     * maxgate x blocks gates each */
    float omega;
    /* replayed-photon numbering: 0 = contiguous, idx*threadphoton +
     * min(idx, oddphotons) + k; 1 = the reference's idx*threadphoton +
     * min(idx, oddphotons - 1) + k (src/mcx_core.cu:1206, :1701, :2021) */
    int replayindex;
} OracleParams;

typedef struct OracleOut {
    double* field;        /* planes*blocks*maxgate*nx*ny*nz FP64 tally, or NULL (planes 2 for RF, blocks =
                             detnum for replaydet -1) */
    float*  field32;      /* twice that: FP32 tally + MAX_ACCUM shadow, or NULL */
    float*  genergy;      /* 2*nthread: escaped, launched (must be pre-zeroed) */
    float*  detdata;      /* maxdetphoton*reclen records, or NULL */
    unsigned int detected;
    unsigned int reclen;
    int errors;           /* "should never happen" exits */
    unsigned long long nstep, ndeposit, nscatter, nlaunch;
    double seconds;       /* wall time of the transport loop */
    unsigned int* seeddata;  /* maxdetphoton*4 words (with saveseed), or NULL */
    unsigned long long* thread_steps;  /* nthread step counts (load balance), or NULL */
} OracleOut;

/* Run the transport loop.  Returns 0 on success, <0 on bad input. */
int mcx_oracle_run(const OracleParams* p, OracleOut* o);

/* glibc rand() stream (TYPE_3 additive feedback) after srand(seed); seed 0
 * and 1 give the default state.  Writes n outputs. */
void mcx_oracle_glibc_rand(unsigned int seed, unsigned int n, int* out);

/* The reference's per-thread seed words (src/mcx_core.cu:3381-3387):
 * words[i] = (r<<16) | (r'<<1) | (r''>>14), three successive rand() draws. */
void mcx_oracle_seedwords(unsigned int seed, unsigned int nwords, unsigned int* words);

/* xorshift128+ floats for lane 0 seeded by the first four seed words
 * (src/mcx_rand_xorshift128p.cu:56-82). */
void mcx_oracle_rng_floats(const unsigned int* seed4, unsigned int n, float* out);

/* The detector mask (src/mcx_utils.c:4214-4340) on a uint32 column-major
 * label volume, in place. */
void mcx_oracle_maskdet(unsigned int* vol, unsigned int nx, unsigned int ny, unsigned int nz,
                        const float* detpos, unsigned int detnum, int svmc);

/* Math entry points for tests. */
float mcx_oracle_expf(float x);
float mcx_oracle_logf(float x);
void  mcx_oracle_sincos2pi(float u, float* s, float* c);
void  mcx_oracle_sincosf(float x, float* s, float* c);

/* The transport loop's own pieces, exposed so tests can pin each one
 * against closed forms (the loop calls exactly these functions):
 * the scattering cos(theta) for g and a uniform u: HG inversion, or
 * isotropic when |g| <= FLT_EPSILON (src/mcx_core.cu:1946-1964);
 * Fresnel reflectance at a voxel face for |cos| of the incidence angle
 * (:2418-2432); the void-entry reflectance used with isspecular
 * (reflectcoeff, :552-570); the refraction of a unit vector across `face`
 * (transmit, :523-537). */
float mcx_oracle_hg_costheta(float g, float u);
float mcx_oracle_fresnel_face(float n1, float n2, float cosi);
float mcx_oracle_reflectcoeff(float n1, float n2, float cosi);
void  mcx_oracle_transmit(float n1, float n2, int face, const float* vin, float* vout);

/* Straight-line voxel stepping of the loop (hitgrid + face crossing,
 * src/mcx_core.cu:493-507, :2127-2132) from p0 along unit v0: writes the
 * voxel entered (3 ints) and the step length of each crossing until the ray
 * leaves the nx*ny*nz grid or maxn steps.  Returns the step count. */
int mcx_oracle_trace(const float* p0, const float* v0, unsigned int nx, unsigned int ny, unsigned int nz,
                     unsigned int maxn, int* ids, float* lens);

/* n successive photon launches of virtual thread 0 (launchnewphoton,
 * src/mcx_core.cu:1089-1717) for the source in P, its RNG seeded from
 * P->seeds[0..3]: out[8i..8i+7] = position x,y,z, weight, direction
 * x,y,z, w.  Returns n, or < 0 when a launch fails. */
int mcx_oracle_launch_samples(const OracleParams* P, unsigned int n, float* out);

/* Planted mistakes, for the tests that prove the pins catch a mistake:
 *  0 none;  1 HG cos(theta) with the wrong sign;  2 HG square dropped;
 *  3 face Fresnel p-term equal to the s-term (n1, n2 transposed);
 *  4 hitgrid always takes the upper face;  5 Snell ratio inverted in
 *  transmit;  6 reflectcoeff with n1, n2 swapped;  11 fourierx second
 *  edge reversed;  12 fourierx2d phases swapped;  13 zgaussian angle
 *  scale sqrt(E) instead of sqrt(2E);  14 line fan one-sided;  15 slit
 *  broadening widths swapped;  16 hyperboloid focus on the wrong side;
 *  17 pattern3d x/y indices transposed;  18 arcsine zenith over [0, pi/2];
 *  19 ring radius uniform instead of area-uniform;  20 RF phasor with sin
 *  and cos exchanged;  21 sincos(x) reduction without its last pi/2 part.
 * Global to the library: set between runs only. */
void  mcx_oracle_set_mutation(int id);

#ifdef __cplusplus
}
#endif

#endif
