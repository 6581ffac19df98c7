/*
 * mcx_abi.h -- C ABI of the MI355X photon-transport engine.
 *
 * This header is the drop-in boundary.  Every struct below keeps the byte
 * layout of the reference's data contract so that an existing caller (the
 * `mcx` CLI main, the pmcx binding, a MATLAB gateway) can be relinked
 * against this library unchanged:
 *
 *   Medium     src/mcx_utils.h:83-88
 *   POLMedium  src/mcx_utils.h:94-100
 *   ExtraSrc   src/mcx_utils.h:102-107
 *   History    src/mcx_utils.h:115-132   (64 bytes: the .mch header)
 *   Replay     src/mcx_utils.h:138-143
 *   GPUInfo    src/mcx_utils.h:149-166
 *   Config     src/mcx_utils.h:172-285
 *
 * The reference builds these from CUDA's vector types.  This library has
 * its own POD vector types with the same size and alignment: 4-component
 * types are 16-byte aligned, 3-component types are 4-byte aligned, and
 * 2-component types are 8-byte aligned.  They carry the same member names
 * (x, y, z, w), so caller source code that reads cfg->srcpos.x compiles
 * unchanged.
 *
 * Plain C: no HIP or torch types appear in any signature.
 */
#ifndef MCX_AMD_ABI_H
#define MCX_AMD_ABI_H

#include <stddef.h>
#include <stdio.h>
#include <float.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- POD vector types (layout-compatible with the reference's) ------- */

typedef struct mcx_float2 { float x, y; } __attribute__((aligned(8))) mcx_float2;
typedef struct mcx_float3 { float x, y, z; } mcx_float3;
typedef struct mcx_float4 { float x, y, z, w; } __attribute__((aligned(16))) mcx_float4;
typedef struct mcx_uint2 { unsigned int x, y; } __attribute__((aligned(8))) mcx_uint2;
typedef struct mcx_uint3 { unsigned int x, y, z; } mcx_uint3;
typedef struct mcx_uint4 { unsigned int x, y, z, w; } __attribute__((aligned(16))) mcx_uint4;

/* ---- constants (src/mcx_const.h:38-128, src/mcx_utils.h:42-60) -------- */

#define MCX_VERSION          "v2025.10-mi355x"
#define MCX_EPS              FLT_EPSILON
#define MCX_VERY_BIG         (1.f / FLT_EPSILON)
#define MCX_MAX_FULL_PATH    2048
#define MCX_MAX_PATH_LENGTH  1024
#define MCX_MAX_SESSION_LENGTH 256
#define MCX_MAX_DEVICE       256
#define MCX_MAX_LANG_ID      32
#define MCX_CUDA_ERROR_LAUNCH_FAILED 719

#define MCX_ONE_PI           3.1415926535897932f
#define MCX_TWO_PI           6.28318530717959f
#define MCX_C0               299792458000.f          /* mm/s */
#define MCX_R_C0             3.335640951981520e-12f  /* s/mm */
#define MCX_JUST_BELOW_ONE   0.9998f
#define MCX_FILL_MAXDETPHOTON 3
#define MCX_OUTSIDE_VOLUME_MIN 0xFFFFFFFFu
#define MCX_OUTSIDE_VOLUME_MAX 0x7FFFFFFFu
#define MCX_MAX_PROP_AND_DETECTORS 4000
#define MCX_SEED_FROM_FILE   (-999)
#define MCX_DET_MASK         0x80000000u
#define MCX_MED_MASK         0x7FFFFFFFu
#define MCX_MAX_ACCUM        1000.f                  /* src/mcx_core.h:53 */
#define MCX_ROULETTE_SIZE    10.f                    /* src/mcx_core.h:55 */

#include "mcx_media.h"

#define MCX_DEBUG_RNG        1
#define MCX_DEBUG_MOVE       2
#define MCX_DEBUG_PROGRESS   4
#define MCX_DEBUG_MOVE_ONLY  8

/* source types (src/mcx_const.h:84-101) */
enum {
    MCX_SRC_PENCIL = 0, MCX_SRC_ISOTROPIC = 1, MCX_SRC_CONE = 2, MCX_SRC_GAUSSIAN = 3,
    MCX_SRC_PLANAR = 4, MCX_SRC_PATTERN = 5, MCX_SRC_FOURIER = 6, MCX_SRC_ARCSINE = 7,
    MCX_SRC_DISK = 8, MCX_SRC_FOURIERX = 9, MCX_SRC_FOURIERX2D = 10, MCX_SRC_ZGAUSSIAN = 11,
    MCX_SRC_LINE = 12, MCX_SRC_SLIT = 13, MCX_SRC_PENCILARRAY = 14, MCX_SRC_PATTERN3D = 15,
    MCX_SRC_HYPERBOLOID_GAUSSIAN = 16, MCX_SRC_RING = 17
};

/* detected-photon record bits (src/mcx_const.h:103-110) */
#define MCX_SAVE_DETID(a)  ((a)      & 0x1)
#define MCX_SAVE_NSCAT(a)  ((a) >> 1 & 0x1)
#define MCX_SAVE_PPATH(a)  ((a) >> 2 & 0x1)
#define MCX_SAVE_MOM(a)    ((a) >> 3 & 0x1)
#define MCX_SAVE_PEXIT(a)  ((a) >> 4 & 0x1)
#define MCX_SAVE_VEXIT(a)  ((a) >> 5 & 0x1)
#define MCX_SAVE_W0(a)     ((a) >> 6 & 0x1)
#define MCX_SAVE_IQUV(a)   ((a) >> 7 & 0x1)

/* enums (src/mcx_utils.h:62-70) */
enum TOutputType { otFlux, otFluence, otEnergy, otJacobian, otWP, otDCS, otRF, otL, otRFmus, otWLTOF, otWPTOF };
enum TMCXParent  { mpStandalone, mpMATLAB, mpPython };
enum TOutputFormat { ofMC2, ofNifti, ofAnalyze, ofUBJSON, ofTX3, ofJNifti, ofBJNifti };
enum TBoundary   { bcUnknown, bcReflect, bcAbsorb, bcMirror, bcCyclic };

/* ---- ABI structs ------------------------------------------------------- */

typedef struct MCXMedium {              /* src/mcx_utils.h:83-88 */
    float mua, mus, g, n;
} Medium;

typedef struct MCXPolarizeMedium {      /* src/mcx_utils.h:94-100 */
    float mua, r, rho, nsph, nmed;
} POLMedium;

typedef struct MCXExtraSource {         /* src/mcx_utils.h:102-107 */
    mcx_float4 srcpos, srcdir, srcparam1, srcparam2;
} ExtraSrc;

typedef struct MCXHistoryHeader {       /* src/mcx_utils.h:115-132 */
    char magic[4];
    unsigned int version, maxmedia, detnum, colcount, totalphoton, detected, savedphoton;
    float unitinmm;
    unsigned int seedbyte;
    float normalizer;
    int respin;
    unsigned int srcnum, savedetflag, totalsource;
    int reserved[1];
} History;

typedef struct PhotonReplay {           /* src/mcx_utils.h:138-143 */
    int* detid;
    void* seed;
    float* weight;
    float* tof;
} Replay;

typedef struct MCXGPUInfo {             /* src/mcx_utils.h:149-166 */
    char name[MCX_MAX_SESSION_LENGTH];
    int id, devcount, major, minor;
    size_t globalmem, constmem, sharedmem;
    int regcount, clock, sm, core, autoblock, autothread;
    unsigned int maxgate;
    int maxmpthread;
} GPUInfo;

typedef struct MCXConfig {              /* src/mcx_utils.h:172-285 */
    size_t nphoton;
    unsigned int nblocksize;
    unsigned int nthread;
    int seed;
    mcx_float4 srcpos;
    mcx_float4 srcdir;
    mcx_float4 srciquv;
    float tstart, tstep, tend;
    mcx_float3 steps;
    mcx_uint3 dim, crop0, crop1;
    unsigned int medianum, polmedianum, detnum, maxdetphoton;
    float detradius, sradius;
    Medium* prop;
    POLMedium* polprop;
    mcx_float4* detpos;
    mcx_float4* smatrix;
    unsigned int maxgate;
    int respin, printnum, gpuid;
    unsigned int* vol;
    char session[MCX_MAX_SESSION_LENGTH];
    char isrowmajor, isreflect, isref3, isrefint, isnormalized, issavedet, issave2pt,
         isgpuinfo, isspecular, issrcfrom0, isdumpmask, autopilot, issaveseed, issaveexit,
         issaveref, ismomentum, istrajstokes, isdumpjson, internalsrc;
    int zipid;
    char srctype, outputtype, outputformat, faststep;
    float minenergy, unitinmm, omega, lambda;
    FILE* flog;
    History his;
    float* exportfield;
    float* exportdetected;
    unsigned long int detectedcount;
    char rootpath[MCX_MAX_PATH_LENGTH];
    char* shapedata;
    char* extrajson;
    int maxvoidstep, voidtime;
    mcx_float4 srcparam1, srcparam2;
    unsigned int srcnum;
    float* srcpattern;
    Replay replay;
    void* seeddata;
    int replaydet;
    char seedfile[MCX_MAX_PATH_LENGTH];
    char jsonfile[MCX_MAX_PATH_LENGTH];
    unsigned int debuglevel, savedetflag;
    char deviceid[MCX_MAX_DEVICE];
    float workload[MCX_MAX_DEVICE];
    int parentid;
    unsigned int runtime;
    char langid[MCX_MAX_LANG_ID];
    double energytot, energyabs, energyesc;
    float normalizer;
    unsigned int maxjumpdebug, debugdatalen, gscatter;
    float* exportdebugdata;
    unsigned int mediabyte;
    float* dx;
    float* dy;
    float* dz;
    char bc[13];
    unsigned int nphase;
    float* invcdf;
    unsigned int nangle;
    float* angleinvcdf;
    int srcid;
    unsigned int extrasrclen;
    ExtraSrc* srcdata;
} Config;

/* ---- exported entry points ---------------------------------------------
 * Each replaces the reference function of the same name and keeps its
 * argument meaning.  Errors follow mcx_error's convention: a standalone
 * caller (cfg->parentid == mpStandalone) gets "MCX ERROR(id):..." on stderr
 * and exit(id); an embedded caller (mpPython, mpMATLAB) gets the call
 * abandoned and the error readable through mcx_last_error().
 */

/* src/mcx_core.h:200, src/mcx_core.cu:2740 -- run the simulation on the
 * device selected by cfg->deviceid; fills cfg->exportfield (accumulated,
 * caller zeroes it), cfg->exportdetected, cfg->detectedcount, cfg->his,
 * cfg->energytot/energyesc/energyabs, cfg->normalizer, cfg->runtime (ms). */
void mcx_run_simulation(Config* cfg, GPUInfo* gpu);

/* src/mcx_core.h:201, src/mcx_core.cu:2621 -- enumerate HIP devices,
 * compact cfg->deviceid to 1-based active ids, return the active count. */
int  mcx_list_gpu(Config* cfg, GPUInfo** info);

/* src/mcx_utils.h:296, src/mcx_utils.c:248 -- defaults. */
void mcx_initcfg(Config* cfg);
/* src/mcx_utils.h:297, src/mcx_utils.c:411 -- free what the library owns. */
void mcx_clearcfg(Config* cfg);
/* src/mcx_utils.h:298, src/mcx_utils.c:1525 -- normalise and mask. */
void mcx_preprocess(Config* cfg);
/* src/mcx_utils.h:311, src/mcx_utils.c:4214 -- DET_MASK on detector voxels. */
void mcx_maskdet(Config* cfg);
/* src/mcx_utils.h:338, src/mcx_utils.c:3748 -- embedded-mode validation
 * (1-based -> 0-based positions, gates, then mcx_preprocess). */
void mcx_validatecfg(Config* cfg, float* detps, int dimdetps[2], int seedbyte);
/* src/mcx_utils.h:324, src/mcx_utils.c -- free a GPUInfo list. */
void mcx_cleargpuinfo(GPUInfo** gpuinfo);
/* src/mcx_utils.h:292, src/mcx_utils.c:1314 -- report an error. */
void mcx_error(const int id, const char* msg, const char* file, const int linenum);
/* src/mcx_utils.h:327 -- flush the log stream. */
void mcx_flush(Config* cfg);
/* src/mcx_utils.h:304, src/mcx_utils.c:1227 -- field *= scale. */
void mcx_normalize(float field[], float scale, size_t fieldlen, int option, int pidx, int srcnum);
/* src/mcx_utils.h:340, src/mcx_utils.c:5490 -- two floats packed as the
 * half-precision word of the muamus_float / asgn_float media formats, with
 * the reference encoder's exact bit rules. */
int  mcx_float2half2(float input[2]);
/* src/mcx_utils.h:300, src/mcx_utils.c:1418 -- load a JSON input file
 * (or, if fname starts with '{', inline JSON text) into cfg and prepare
 * the domain (shapes, preprocess). */
void mcx_readconfig(char* fname, Config* cfg);
/* src/mcx_utils.h:328, src/mcx_utils.c:5578 -- parse JSON, run, return 0. */
int  mcx_run_from_json(char* jsonstr);
/* src/mcx_utils.h:332, src/mcx_utils.c:916 -- write the fluence output
 * (.jnii by default, .mc2 raw float32 for outputformat ofMC2). */
void mcx_savedata(float* dat, size_t len, Config* cfg);
/* src/mcx_utils.h:326, src/mcx_utils.c:982 -- write detected photons
 * (.mch: History header + float32 rows + optional seeds, or _detp.jdat). */
void mcx_savedetphoton(float* ppath, void* seeds, int count, int seedbyte, Config* cfg);

/* Error channel for embedded callers (no reference counterpart: the
 * reference throws a C++ `const char*` across this boundary in container
 * builds, src/pmcx.cpp:1369-1372, which a C caller cannot catch).
 * Returns the last error id (0 if none) and copies its message. */
int  mcx_last_error(char* msg, int msglen);
void mcx_clear_error(void);

/* RNG known-answer entry (src/mcx_core.cu:1731-1742, :3056-3114, the -D R
 * path): fill out[0..n) with n draws of lane 0's xorshift128+ stream seeded
 * from four glibc-rand() words taken at the state srand(seed) leaves
 * (seed == 0 keeps the default state, as when no srand() preceded).
 * Runs on the GPU.  Returns 0 on success. */
int  mcx_rng_test(unsigned int seed, float* out, unsigned int n);

#ifdef __cplusplus
}
#endif

#endif /* MCX_AMD_ABI_H */
