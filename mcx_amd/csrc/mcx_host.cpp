/*
 * mcx_host.cpp -- launch/reduction layer: mcx_run_simulation, mcx_list_gpu.
 *
 * Follows the reference's host flow (src/mcx_core.cu:2740-4110): derive the
 * launch constants, seed every virtual thread from the glibc rand() stream,
 * loop over time-gate groups and respins, launch the transport kernel, then
 * merge the tally, sum the energies in FP64 and normalise.  MI355X
 * differences:
 *  - one HIP stream, hipEvents for kernel time, steady_clock around
 *    launch+sync for the reference's `runtime` definition;
 *  - the tally halves (shadow or hi/lo) are merged on the GPU by a
 *    vectorised kernel instead of a serial host loop (:3687-3696);
 *  - the label volume is shipped as uint8 (bit 7 = detector mask) when the
 *    labels fit, uint32 otherwise;
 *  - autopilot sizes the grid from the kernel's register count so every
 *    SIMD of every CU is filled once (256 CUs x 4 SIMDs x waves/SIMD).
 */
#include <hip/hip_runtime_api.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include <algorithm>
#include <cmath>
#include <chrono>
#include <mutex>
#include <vector>

#include "../../include/mcx_amd_ext.h"
#include "mcx_device.h"

/* dkp: the launch parameters in device memory */
extern "C" hipError_t mcxk_launch_transport(const KParam* dkp, int ispencil, int isreflect, int tally, int media32,
                                            unsigned int grid, unsigned int block, size_t shmem, hipStream_t stream);
extern "C" hipError_t mcxk_kernel_attrs(int ispencil, int isreflect, int tally, int media32, int* numregs, int* maxthreads);
extern "C" hipError_t mcxk_launch_rngtest(float* out, const uint32_t* seeds, uint32_t n, hipStream_t stream);
extern "C" hipError_t mcxk_launch_merge(int tally, float* out, const void* a, const void* b, const void* c, size_t n, hipStream_t stream);
extern "C" hipError_t mcxk_launch_replica_cw(float* cw, const double* d, const double* scale, uint32_t nx, uint32_t ny,
                                             uint32_t nz, uint32_t planes, uint32_t nmaps, uint32_t bnx, uint32_t bny,
                                             uint32_t bnz, hipStream_t stream);
extern "C" hipError_t mcxk_launch_merge_blocked(float* out, const double* a, uint32_t nx, uint32_t ny, uint32_t nz, uint32_t bnx,
                                                 uint32_t bny, uint32_t bnz, size_t n, hipStream_t stream);

namespace mcx {

static std::mutex g_statmu;
static double g_stats[18];
static int g_tally_mode = MCX_TALLY_F64;
static int g_counters = 0;
static int g_hotbox = 1;
/* replayed-photon numbering: 0 contiguous, 1 the reference's */
static int g_replayindex = 0;
/* blocked FP64 tally with per-lane LDS blocks (KParam::tblock): 0 off,
 * 1 on, 2 (default) on when the medium's mean scattering coefficient is
 * at most 2 per voxel length: where photons scatter many times within a
 * voxel, deposits are few per step, the atomics do not bound the kernel
 * and the blocks only add instructions */
static int g_tblock = 2;
/* lane refill (autopilot lanes run several virtual threads each,
 * KParam::refill): 0 off, 1 (default) at least kMinStreamPhotons photons per
 * virtual thread and at most kDefaultRefill virtual threads per lane, 2
 * about one photon per virtual thread */
static int g_refill = 1;
/* virtual threads per lane at most under refill (each holds 24 bytes of
 * seeds and energies) */
static const unsigned int kMaxRefill = 32, kDefaultRefill = 8;
/* The first xorshift128+ output of a stream seeded with the host's seed
 * words is not uniform (the words OR overlapping rand() values, so their
 * bits are biased towards 1): its mean is 0.4914 and -log of it averages
 * 1.039, so the first photon of each thread starts with a longer
 * scattering length.  The reference draws the same numbers; the effect on
 * a result shrinks as 1/(photons per thread) (cube60: +0.018/P relative in
 * the absorbed fraction), so refill 1 keeps P >= 64: a shift below 2.8e-4,
 * the precision of the reference's published four-digit answers. */
static const double kMinStreamPhotons = 64.0;
/* a request for K replicas of the next run (mcx_amd_run_replicas) */
struct ReplicaReq {
    int k;
    float* cw;
    double* energy;
    double* stats;
};
static ReplicaReq* g_rep = nullptr;

static bool supported_source(int t) { return t >= MCX_SRC_PENCIL && t <= MCX_SRC_RING; }

/* property-word media formats the kernel decodes (src/mcx_const.h:72-80) */
int media_words(unsigned int mediabyte) {
    return mediabyte >= MCX_MEDIA_ASGN_F2H && mediabyte <= MCX_MEDIA_AS_SHORT;
}

/* 32-bit words per voxel of the uploaded volume */
size_t media_words_per_voxel(unsigned int mediabyte) {
    return (mediabyte == MCX_MEDIA_ASGN_F2H || mediabyte == MCX_MEDIA_2LABEL_SPLIT) ? 2 : 1;
}

/* number of pattern weights of a pattern / pattern3d source */
static size_t pattern_len(const Config* cfg) {
    if (cfg->srctype == MCX_SRC_PATTERN) return (size_t)cfg->srcparam1.w * (size_t)cfg->srcparam2.w;
    if (cfg->srctype == MCX_SRC_PATTERN3D)
        return (size_t)cfg->srcparam1.x * (size_t)cfg->srcparam1.y * (size_t)cfg->srcparam1.z;
    return 0;
}

void check_supported(const Config* cfg) {
    if (cfg->polmedianum > 0) MCX_THROW(-1, "polarized photon simulation is not supported by this engine");
    if (cfg->mediabyte > 4 && !media_words(cfg->mediabyte))
        MCX_THROW(-1, "the specified media format is not supported by this engine");
    const bool replay = cfg->seed == MCX_SEED_FROM_FILE;
    if (replay && (!cfg->replay.seed || !cfg->replay.weight || !cfg->replay.tof || !cfg->replay.detid))
        MCX_THROW(-6, "photon replay needs the saved seeds and the detected-photon data (run mcx_validatecfg with detphotons)");
    if (cfg->nphase || cfg->nangle) MCX_THROW(-1, "user-defined phase/angle inverse CDFs are not supported by this engine");
    if (cfg->issaveref > 1) MCX_THROW(-1, "per-voxel exit partial paths (issaveref > 1) are not supported by this engine");
    if (cfg->extrasrclen) MCX_THROW(-1, "multiple sources are not supported by this engine");
    if (cfg->debuglevel & (MCX_DEBUG_MOVE | MCX_DEBUG_MOVE_ONLY))
        MCX_THROW(-1, "trajectory output (-D M) is not supported by this engine");
    if (!supported_source(cfg->srctype)) MCX_THROW(-1, "the specified source type is not supported by this engine");
    if (pattern_len(cfg) && cfg->srcnum > 1)
        MCX_THROW(-1, "multiple simultaneous source patterns (srcnum > 1) are not supported by this engine");
    if ((cfg->srctype == MCX_SRC_PATTERN || cfg->srctype == MCX_SRC_PATTERN3D) && (!cfg->srcpattern || pattern_len(cfg) == 0))
        MCX_THROW(-4, "the 'srcpattern' field can not be empty when your 'srctype' is 'pattern'");
    const bool replaytype = cfg->outputtype == otJacobian || cfg->outputtype == otWP || cfg->outputtype == otDCS ||
                            cfg->outputtype == otRF || cfg->outputtype == otRFmus || cfg->outputtype == otWLTOF ||
                            cfg->outputtype == otWPTOF;
    if (replaytype && !replay) MCX_THROW(-6, "Jacobian output is only valid in the reply mode. Please define cfg.seed");
    if (cfg->outputtype != otFlux && cfg->outputtype != otFluence && cfg->outputtype != otEnergy && cfg->outputtype != otL &&
        !replaytype)
        MCX_THROW(-1, "the specified output type is not supported by this engine");
}

/* waves per SIMD for a kernel using `vgprs` architectural VGPRs (gfx950:
 * 512 per lane per SIMD, granule 8, at most 8 waves) */
int counters_enabled() { return g_counters; }

int waves_per_simd(int vgprs) {
    if (vgprs <= 0) return 8;
    int g = (vgprs + 7) / 8 * 8;
    return std::max(1, std::min(8, 512 / g));
}


/* Largest box of cells that can be summed in LDS without lowering the
 * occupancy: c cells over three axes of extents d, each box edge at most
 * its axis' extent, the shortest axis first so the other two can take
 * what it cannot. */
static void box_dims(const uint32_t d[3], uint32_t c, uint32_t b[3]) {
    int ord[3] = {0, 1, 2};
    std::sort(ord, ord + 3, [&](int a, int q) { return d[a] < d[q]; });
    uint32_t left = c;
    for (int k = 0; k < 3; k++) {
        const int ax = ord[k];
        const double root = k == 0 ? std::cbrt((double)left) : (k == 1 ? std::sqrt((double)left) : (double)left);
        b[ax] = std::max(1u, std::min(d[ax], (uint32_t)std::floor(root + 1e-9)));
        left = std::max(1u, left / b[ax]);
    }
}

/* Geometry of a hot-cell box of at most `cells` cells (KParam::hot_*):
 * gate 0 of the time window, voxels around the source's voxel, reaching
 * into the medium along the launch direction, clamped to the grid.  Sets
 * hot_cells = 0 when fewer than 64 cells fit. */
void hot_box_geometry(const Config* cfg, uint32_t cells, KParam& kp) {
    kp.hot_cells = 0;
    if (cells < 64 || kp.maxgate == 0 || kp.nvox == 0) return;
    const uint32_t d[3] = {kp.nx, kp.ny, kp.nz};
    uint32_t b[3];
    box_dims(d, cells, b);
    if (b[0] * b[1] * b[2] < 64) return;
    const float* src = &cfg->srcpos.x;
    const float* dir = &cfg->srcdir.x;
    uint32_t o[3];
    for (int k = 0; k < 3; k++) {
        const double c = std::floor((double)src[k]) + (double)dir[k] * (0.5 * b[k] - 1.0);
        const double lo = std::floor(c - 0.5 * b[k] + 0.5);
        o[k] = (uint32_t)std::max(0.0, std::min((double)(d[k] - b[k]), lo));
    }
    kp.hot_cells = b[0] * b[1] * b[2];
    kp.hot_g = 1;
    kp.hot_x0 = o[0];
    kp.hot_y0 = o[1];
    kp.hot_z0 = o[2];
    kp.hot_bx = b[0];
    kp.hot_by = b[1];
    kp.hot_bz = b[2];
    div_magic(kp.dimxy, &kp.hot_mxy, &kp.hot_shxy);
    div_magic(kp.nx, &kp.hot_mx, &kp.hot_shx);
}

/* LDS bytes each of `groups` co-resident workgroups (or waves) per CU can
 * take: the CU's LDS shared evenly, at most `cap`, in 512-byte granules */
size_t lds_share(int device, unsigned int groups, size_t cap) {
    int lds_cu = 0, lds_blk = 0;
    if (hipDeviceGetAttribute(&lds_cu, hipDeviceAttributeMaxSharedMemoryPerMultiprocessor, device) != hipSuccess ||
        hipDeviceGetAttribute(&lds_blk, hipDeviceAttributeMaxSharedMemoryPerBlock, device) != hipSuccess || lds_cu <= 0 ||
        lds_blk <= 0)
        return 0;
    size_t per = std::min<size_t>({cap, (size_t)lds_blk, (size_t)lds_cu / std::max(1u, groups)});
    return per & ~(size_t)511;
}

/* The hot-cell box of a single-simulation launch: as large as the LDS a
 * workgroup can take at the kernel's full occupancy allows (at most the
 * device's per-workgroup LDS in all).  Returns the workgroup's total
 * dynamic LDS bytes. */
static size_t set_hot_box(const Config* cfg, int tally, unsigned int block, int numregs, size_t shmem, KParam& kp) {
    kp.hot_cells = 0;
    if (!g_hotbox || tally == MCX_TALLY_F32_SHADOW || kp.maxgate == 0 || kp.nvox == 0) return shmem;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return shmem;
    const unsigned int waves_cu = 4u * (unsigned int)waves_per_simd(numregs);
    const unsigned int wgs = std::max(1u, waves_cu / std::max(1u, block / 64u));
    const size_t per = lds_share(dev, wgs, SIZE_MAX);
    const size_t off = (shmem + 7) & ~(size_t)7;
    if (per <= off + 64 * sizeof(double)) return shmem;
    hot_box_geometry(cfg, (uint32_t)((per - off) / sizeof(double)), kp);
    if (!kp.hot_cells) return shmem;
    kp.hot_off = (uint32_t)off;
    return off + (size_t)kp.hot_cells * sizeof(double);
}

/* launch constants of one simulation (buffers, work split and time
 * window are set by the caller) */
void fill_kparam(const Config* cfg, unsigned int maxgate, unsigned int is2d, KParam& kp) {
    memset(&kp, 0, sizeof(kp));
    const size_t nvox = (size_t)cfg->dim.x * cfg->dim.y * cfg->dim.z;
    const unsigned int partialdata = det_partialdata(cfg);
    const unsigned int hostdetreclen = det_reclen(cfg);
    const unsigned int ntable = cfg->medianum + cfg->detnum;
    kp.nx = cfg->dim.x;
    kp.ny = cfg->dim.y;
    kp.nz = cfg->dim.z;
    kp.dimxy = cfg->dim.x * cfg->dim.y;
    kp.nvox = (uint32_t)nvox;
    kp.maxidx_x = (float)cfg->dim.x;
    kp.maxidx_y = (float)cfg->dim.y;
    kp.maxidx_z = (float)cfg->dim.z;
    memcpy(kp.srcpos, &cfg->srcpos, 16);
    memcpy(kp.srcdir, &cfg->srcdir, 16);
    memcpy(kp.srcparam1, &cfg->srcparam1, 16);
    memcpy(kp.srcparam2, &cfg->srcparam2, 16);
    kp.srctype = cfg->srctype;
    kp.Rtstep = 1.f / cfg->tstep;
    kp.oneoverc0 = MCX_R_C0 * cfg->unitinmm;
    kp.minaccumtime = 1.f * MCX_R_C0 * cfg->unitinmm;
    kp.maxgate = maxgate;
    kp.minenergy = cfg->minenergy;
    kp.doreflect = (uint32_t)cfg->isreflect;
    kp.isspecular = cfg->isspecular > 0;
    kp.voidtime = (uint32_t)cfg->voidtime;
    kp.maxvoidstep = (uint32_t)cfg->maxvoidstep;
    kp.gscatter = cfg->gscatter;
    kp.outputtype = (uint32_t)cfg->outputtype;
    kp.save2pt = (uint32_t)cfg->issave2pt;
    kp.is2d = is2d;
    kp.bcface = kp.bcdet = 0;
    for (int i = 0; i < 6; i++) {
        kp.bcface |= ((uint32_t)(unsigned char)cfg->bc[i] & 0xFu) << (4 * i);
        kp.bcdet |= (uint32_t)(cfg->bc[6 + i] != 0) << i;
    }
    kp.savedet = (uint32_t)cfg->issavedet;
    kp.savedetflag = cfg->savedetflag;
    kp.maxdetphoton = cfg->maxdetphoton;
    kp.reclen = hostdetreclen;
    kp.partialdata = partialdata;
    kp.maxmedia = cfg->medianum - 1;
    kp.detnum = cfg->detnum;
    kp.medianum = cfg->medianum;
    kp.mediaformat = cfg->mediabyte;
    kp.ntable = ntable;
}

static void rng_test_path(Config* cfg, GPUInfo* gpu, int gpuid, size_t fieldlen) {
    (void)gpu;
    (void)gpuid;
    /* no srand() precedes this path in the reference: glibc default state */
    GlibcRand r(1);
    uint32_t seeds[4];
    r.seed_words(seeds, 4);
    Stream st;
    DevMem dseed, dout;
    dseed.alloc(sizeof(seeds));
    dout.alloc(fieldlen * sizeof(float));
    HIP_CHECK(hipMemcpyAsync(dseed.p, seeds, sizeof(seeds), hipMemcpyHostToDevice, st.s));
    HIP_CHECK(mcxk_launch_rngtest(dout.as<float>(), dseed.as<uint32_t>(), (uint32_t)fieldlen, st.s));
    if (!cfg->exportfield) cfg->exportfield = (float*)calloc(fieldlen, sizeof(float));
    HIP_CHECK(hipMemcpyAsync(cfg->exportfield, dout.p, fieldlen * sizeof(float), hipMemcpyDeviceToHost, st.s));
    HIP_CHECK(hipStreamSynchronize(st.s));
    if (cfg->flog) fprintf(cfg->flog, "generated %zu random numbers\n", fieldlen);
}

/* After a replica launch: launched/escaped energy per replica (lanes
 * idx & (K-1)), each replica's flux normalisation (src/mcx_core.cu:3823-3834)
 * times tstep, and the time-integrated maps reduced on the device.  stats:
 * [0] clear ms, [1] transport kernel ms, [2] reduction ms, [3] launch+sync
 * wall ms of the kernel. */
static void finish_replicas(Config* cfg, ReplicaReq& R, DevMem& dgenergy, DevMem& dfield64, unsigned int nthread, size_t nvox,
                            uint32_t planes, uint32_t bnx, uint32_t bny, uint32_t bnz, EventPair& ev, EventPair& evc,
                            double wall_ms, hipStream_t s) {
    const unsigned int K = (unsigned int)R.k;
    std::vector<float> energy((size_t)nthread * 2);
    HIP_CHECK(hipMemcpyAsync(energy.data(), dgenergy.p, energy.size() * sizeof(float), hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipStreamSynchronize(s));
    std::vector<double> esc(K, 0.0), tot(K, 0.0), scale(K, 0.0);
    for (unsigned int i = 0; i < nthread; i++) {
        esc[i & (K - 1)] += energy[2 * i];
        tot[i & (K - 1)] += energy[2 * i + 1];
    }
    const float Vvox = cfg->steps.x * cfg->steps.y * cfg->steps.z;
    for (unsigned int m = 0; m < K; m++) {
        const float sc = tot[m] > 0.0 ? (float)(cfg->unitinmm / (tot[m] * Vvox * cfg->tstep)) : 0.f;
        scale[m] = (double)sc * (double)cfg->tstep;
        R.energy[2 * m] = esc[m];
        R.energy[2 * m + 1] = tot[m];
    }
    DevMem dscale, dcw;
    dscale.alloc(K * sizeof(double));
    dcw.alloc((size_t)K * nvox * sizeof(float));
    HIP_CHECK(hipMemcpyAsync(dscale.p, scale.data(), K * sizeof(double), hipMemcpyHostToDevice, s));
    EventPair er;
    HIP_CHECK(hipEventRecord(er.a, s));
    HIP_CHECK(mcxk_launch_replica_cw(dcw.as<float>(), dfield64.as<double>(), dscale.as<double>(), cfg->dim.x, cfg->dim.y,
                                     cfg->dim.z, planes, K, bnx, bny, bnz, s));
    HIP_CHECK(hipEventRecord(er.b, s));
    /* R.cw may be host or device memory (unified addressing) */
    HIP_CHECK(hipMemcpyAsync(R.cw, dcw.p, dcw.bytes, hipMemcpyDefault, s));
    HIP_CHECK(hipStreamSynchronize(s));
    float clear_ms = 0.f, kern_ms = 0.f, red_ms = 0.f;
    HIP_CHECK(hipEventElapsedTime(&clear_ms, evc.a, ev.a));
    HIP_CHECK(hipEventElapsedTime(&kern_ms, ev.a, ev.b));
    HIP_CHECK(hipEventElapsedTime(&red_ms, er.a, er.b));
    R.stats[0] = clear_ms;
    R.stats[1] = kern_ms;
    R.stats[2] = red_ms;
    R.stats[3] = wall_ms;
}

/* Pinned host staging for the per-thread seed words, one per host thread,
 * grown as needed and kept until the process ends (refilled launches draw
 * millions of seed words, whose upload from pageable memory takes several
 * times longer).  Every copy from it is waited for before the next launch
 * writes it, and before it is replaced. */
static uint32_t* pinned_seed_buffer(size_t nwords) {
    thread_local uint32_t* buf = nullptr;
    thread_local size_t cap = 0;
    if (nwords > cap) {
        if (buf) HIP_CHECK(hipHostFree(buf));
        buf = nullptr;
        cap = 0;
        const size_t n = std::max<size_t>(nwords, (size_t)1 << 16);
        HIP_CHECK(hipHostMalloc((void**)&buf, n * sizeof(uint32_t), hipHostMallocDefault));
        cap = n;
    }
    return buf;
}

static void run_impl(Config* cfg, GPUInfo* gpu) {
    check_supported(cfg);
    int threadid = 0;
    if (cfg->deviceid[threadid] == '\0') return;
    const int gpuid = cfg->deviceid[threadid] - 1;
    if (gpuid < 0) MCX_THROW(-1, "GPU ID must be non-zero");
    HIP_CHECK(hipSetDevice(gpuid));

    const unsigned int partialdata = det_partialdata(cfg);
    const unsigned int hostdetreclen = det_reclen(cfg);
    const size_t nvox = (size_t)cfg->dim.x * cfg->dim.y * cfg->dim.z;
    const unsigned int totalgates = (unsigned int)((cfg->tend - cfg->tstart) / cfg->tstep + 0.5);
    unsigned int is2d = (cfg->dim.x == 1 ? 1 : (cfg->dim.y == 1 ? 2 : (cfg->dim.z == 1 ? 3 : 0)));
    if (is2d) is2d = is2d * ((cfg->dim.x > 1) + (cfg->dim.y > 1) + (cfg->dim.z > 1) == 2);
    if (is2d) {
        const float* vec = &cfg->srcdir.x;
        if (fabsf(vec[is2d - 1]) > MCX_EPS)
            MCX_THROW(-1, "input domain is 2D, the initial direction can not have non-zero value in the singular dimension");
    }

    GPUInfo& G = gpu[gpuid];
    if (G.maxgate == 0 && nvox > 0) {
        size_t need = nvox + (size_t)cfg->nthread * 64 + sizeof(float) * (size_t)cfg->maxdetphoton * hostdetreclen + (10u << 20);
        size_t free_b = 0, total_b = 0;
        HIP_CHECK(hipMemGetInfo(&free_b, &total_b));
        size_t mg = free_b > need ? (free_b - need) / (nvox * 2 * sizeof(float)) : 1;
        G.maxgate = (unsigned int)std::min<size_t>(std::max<size_t>(mg, 1), (size_t)totalgates);
    }
    cfg->detectedcount = 0;
    cfg->his.detected = 0;
    cfg->his.respin = cfg->respin;
    cfg->his.colcount = hostdetreclen;
    cfg->energytot = cfg->energyabs = cfg->energyesc = 0.0;
    cfg->runtime = 0;

    const int tally = g_rep ? MCX_TALLY_F64 : g_tally_mode;
    const int ispencil = (cfg->srctype == MCX_SRC_PENCIL);
    int isref = cfg->isreflect;
    for (int i = 0; i < 6; i++)
        if (cfg->bc[i] == bcReflect || cfg->bc[i] == bcMirror) isref = 1;

    /* media: uint8 with the detector bit in bit 7 when labels fit, uint32
     * labels, or the 32-bit property words of the non-label formats */
    int media32 = cfg->mediabyte == MCX_MEDIA_2LABEL_SPLIT ? 3 : 2;
    double musgrid = 0.0; /* mean mus x voxel length over tissue voxels */
    if (cfg->mediabyte <= 4) {
        unsigned int maxlabel = 0;
        std::vector<size_t> hist(cfg->medianum, 0);
        for (size_t i = 0; i < nvox; i++) {
            const unsigned int l = cfg->vol[i] & MCX_MED_MASK;
            maxlabel = std::max(maxlabel, l);
            if (l < cfg->medianum) hist[l]++;
        }
        media32 = (maxlabel > 127);
        size_t nt = 0;
        for (unsigned int l = 1; l < cfg->medianum; l++) {
            musgrid += (double)hist[l] * cfg->prop[l].mus * cfg->unitinmm;
            nt += hist[l];
        }
        musgrid = nt ? musgrid / (double)nt : 0.0;
    }

    int numregs = 0, maxthreads = 0;
    HIP_CHECK(mcxk_kernel_attrs(ispencil, isref > 0, tally, media32, &numregs, &maxthreads));
    if (!cfg->autopilot) {
        G.autothread = (int)cfg->nthread;
        G.autoblock = (int)cfg->nblocksize;
        if (cfg->maxgate == 0 || cfg->maxgate > totalgates) cfg->maxgate = totalgates;
        G.maxgate = cfg->maxgate;
    } else {
        G.autoblock = 256;
        G.autothread = G.sm * 4 * waves_per_simd(numregs) * 64;
        if (cfg->maxgate && cfg->maxgate < G.maxgate) G.maxgate = cfg->maxgate;
    }
    if (G.autoblock <= 0 || G.autoblock > 1024 || G.autoblock > maxthreads) G.autoblock = std::min(256, maxthreads);
    if (G.autothread % G.autoblock) G.autothread = (G.autothread / G.autoblock) * G.autoblock;
    if (G.autothread <= 0) MCX_THROW(-1, "thread number must be at least one block");
    const unsigned int maxgate = G.maxgate;
    if (totalgates > maxgate && cfg->isnormalized) {
        if (cfg->flog) fprintf(cfg->flog, "WARNING: GPU memory can not hold all time gates, disabling normalization to allow multiple runs\n");
        cfg->isnormalized = 0;
    }

    /* per-detector replay tallies: one block of gates per detector */
    const bool replay = cfg->seed == MCX_SEED_FROM_FILE;
    const unsigned int blocks = (replay && cfg->replaydet == -1) ? std::max(1u, cfg->detnum) : 1u;
    if (replay && totalgates > maxgate) MCX_THROW(-1, "photon replay needs all time gates in one pass (reduce the gate count)");
    /* RF replay: an imaginary plane follows the real one (src/mcx_core.cu:2909-2911) */
    const bool rf = replay && (cfg->outputtype == otRF || cfg->outputtype == otRFmus);
    /* replicas return time-integrated maps only: every gate of a pass is
     * tallied into one plane (the gate windows still end the photons) */
    const unsigned int tgates = g_rep ? 1u : maxgate;
    const size_t planelen = nvox * tgates * blocks;
    const size_t fieldlen = planelen * (rf ? 2 : 1);
    /* replicas: K independent runs in one launch, one time-integrated plane each */
    const unsigned int nrep = g_rep ? (unsigned int)g_rep->k : 1u;
    if (g_rep) {
        if (replay || cfg->issavedet) MCX_THROW(-1, "replicas are forward runs without detected photons");
        if (totalgates > maxgate || abs(cfg->respin) != 1) MCX_THROW(-1, "replicas need every time gate in one pass and no respin");
        if (cfg->outputtype != otFlux) MCX_THROW(-1, "replicas output the time-integrated flux");
    }
    const size_t tallylen = fieldlen * nrep;
    if (cfg->exportfield == NULL && cfg->issave2pt && !g_rep) cfg->exportfield = (float*)calloc(fieldlen, sizeof(float));
    if (cfg->exportdetected == NULL) cfg->exportdetected = (float*)malloc(std::max<size_t>(1, (size_t)hostdetreclen * cfg->maxdetphoton) * sizeof(float));

    if (cfg->debuglevel & MCX_DEBUG_RNG) {
        rng_test_path(cfg, gpu, gpuid, fieldlen);
        return;
    }

    /* work split (src/mcx_core.cu:2980-3022): one device takes all photons */
    const size_t gpuphoton = cfg->nphoton * nrep;
    if (gpuphoton == 0) return;
    /* Autopilot launches one lane per resident SIMD slot; with refill those
     * lanes run R times as many virtual threads, each lane taking the next
     * unstarted one when its current one is done, so lanes stay busy until
     * the last virtual threads start instead of waiting on the slowest lane
     * of their wave.  R = photons per lane / kMinStreamPhotons, at most
     * kDefaultRefill (refill 1), or photons per lane, at most kMaxRefill
     * (refill 2: about one photon per virtual thread, the shortest tail) */
    const unsigned int nlanes = (unsigned int)G.autothread;
    unsigned int refill = 1;
    if (cfg->autopilot && g_refill && cfg->respin >= 1) {
        const double ppl = (double)gpuphoton / (double)nlanes;
        const double r = g_refill == 2 ? std::min((double)kMaxRefill, std::floor(ppl + 0.5))
                                       : std::min((double)kDefaultRefill, std::floor(ppl / kMinStreamPhotons));
        refill = (unsigned int)std::max(1.0, r);
    }
    G.autothread = (int)(nlanes * refill);
    int threadphoton, oddphotons;
    const unsigned int nthread = (unsigned int)G.autothread;
    if (cfg->respin >= 1) {
        threadphoton = (int)(gpuphoton / nthread);
        oddphotons = (int)(gpuphoton - (size_t)threadphoton * nthread);
    } else {
        threadphoton = (int)(gpuphoton / nthread / (size_t)(-cfg->respin));
        oddphotons = (int)(gpuphoton / (size_t)(-cfg->respin) - (size_t)threadphoton * nthread);
    }
    /* the reference's replay numbering (src/mcx_core.cu:1206) starts thread
     * 0 at photon -1 when every thread takes the same count */
    if (replay && g_replayindex && oddphotons == 0 && threadphoton > 0)
        MCX_THROW(-6, "the reference's replay photon numbering reads before the first saved photon when the photon count "
                      "is a multiple of the thread count; change the thread count or use the contiguous numbering");

    Stream st;
    DevMem dkp, dmedia, dprop, dfield, dfield2, dfieldk, dfield64, dgenergy, dseeds, ddet, dcount, dcounters, dmerged;
    DevMem dsaved, drseed, drweight, drtof, drdetid, dpattern;
    if (pattern_len(cfg)) {
        dpattern.alloc(pattern_len(cfg) * sizeof(float));
        HIP_CHECK(hipMemcpyAsync(dpattern.p, cfg->srcpattern, dpattern.bytes, hipMemcpyHostToDevice, st.s));
    }
    std::vector<unsigned char> m8;
    if (media32) {
        const size_t words = nvox * (media32 >= 2 ? media_words_per_voxel(cfg->mediabyte) : 1);
        dmedia.alloc(words * 4);
        HIP_CHECK(hipMemcpyAsync(dmedia.p, cfg->vol, words * 4, hipMemcpyHostToDevice, st.s));
    } else {
        m8.resize(nvox);
        for (size_t i = 0; i < nvox; i++) m8[i] = (unsigned char)((cfg->vol[i] & 0x7Fu) | ((cfg->vol[i] >> 24) & 0x80u));
        dmedia.alloc(nvox);
        HIP_CHECK(hipMemcpyAsync(dmedia.p, m8.data(), nvox, hipMemcpyHostToDevice, st.s));
    }
    const unsigned int ntable = cfg->medianum + cfg->detnum;
    if (ntable > MCX_MAX_PROP_AND_DETECTORS) MCX_THROW(-1, "too many media and detectors");
    std::vector<float> table(4 * ntable);
    memcpy(table.data(), cfg->prop, sizeof(float) * 4 * cfg->medianum);
    if (cfg->detnum) memcpy(table.data() + 4 * cfg->medianum, cfg->detpos, sizeof(float) * 4 * cfg->detnum);
    dprop.alloc(table.size() * sizeof(float));
    HIP_CHECK(hipMemcpyAsync(dprop.p, table.data(), table.size() * sizeof(float), hipMemcpyHostToDevice, st.s));

    const bool savedet = cfg->issavedet != 0;
    /* blocked FP64 tally: 2 x 2 x 2 voxel blocks of every gate plane */
    const uint32_t bnx = (cfg->dim.x + 1) / 2, bny = (cfg->dim.y + 1) / 2, bnz = (cfg->dim.z + 1) / 2;
    const uint64_t nblk = (uint64_t)(tallylen / nvox) * bnz * bny * bnx;
    bool tblock = (g_tblock == 1 || (g_tblock == 2 && musgrid <= 2.0)) && tally == MCX_TALLY_F64 && nblk < 0xFFFFFFFFull;
    if (tally == MCX_TALLY_F64) {
        dfield64.alloc((tblock ? std::max<size_t>(tallylen, (size_t)nblk * 8) : tallylen) * sizeof(double));
    } else if (tally == MCX_TALLY_F64_XCD) {
        dfield64.alloc(fieldlen * MCX_NUM_XCD * sizeof(double));
    } else {
        dfield.alloc(fieldlen * sizeof(float));
        dfield2.alloc(fieldlen * sizeof(float));
        if (tally == MCX_TALLY_F32_HILO) dfieldk.alloc(fieldlen * sizeof(int32_t));
    }
    if (!g_rep) dmerged.alloc(fieldlen * sizeof(float));
    dgenergy.alloc((size_t)nthread * 2 * sizeof(float));
    HIP_CHECK(hipMemsetAsync(dgenergy.p, 0, dgenergy.bytes, st.s));
    dseeds.alloc((size_t)nthread * 4 * sizeof(uint32_t));
    if (savedet) ddet.alloc(std::max<size_t>(1, (size_t)cfg->maxdetphoton * hostdetreclen) * sizeof(float));
    dcount.alloc(4 * sizeof(uint32_t));
    dcounters.alloc(7 * sizeof(uint64_t));
    const bool saveseed = savedet && cfg->issaveseed;
    if (saveseed) dsaved.alloc(std::max<size_t>(1, (size_t)cfg->maxdetphoton) * 16);
    if (replay) {
        /* one saved state, weight, time of flight and detector id per replayed photon */
        drseed.alloc(std::max<size_t>(1, gpuphoton) * 16);
        drweight.alloc(std::max<size_t>(1, gpuphoton) * sizeof(float));
        drtof.alloc(std::max<size_t>(1, gpuphoton) * sizeof(float));
        drdetid.alloc(std::max<size_t>(1, gpuphoton) * sizeof(int));
        HIP_CHECK(hipMemcpyAsync(drseed.p, cfg->replay.seed, gpuphoton * 16, hipMemcpyHostToDevice, st.s));
        HIP_CHECK(hipMemcpyAsync(drweight.p, cfg->replay.weight, gpuphoton * sizeof(float), hipMemcpyHostToDevice, st.s));
        HIP_CHECK(hipMemcpyAsync(drtof.p, cfg->replay.tof, gpuphoton * sizeof(float), hipMemcpyHostToDevice, st.s));
        HIP_CHECK(hipMemcpyAsync(drdetid.p, cfg->replay.detid, gpuphoton * sizeof(int), hipMemcpyHostToDevice, st.s));
    }
    dkp.alloc(sizeof(KParam));
    HIP_CHECK(hipMemsetAsync(dcounters.p, 0, dcounters.bytes, st.s));

    KParam kp;
    fill_kparam(cfg, maxgate, is2d, kp);
    kp.maxgate = tgates;
    kp.threadphoton = threadphoton;
    kp.oddphotons = oddphotons;
    kp.nthread = nthread;
    kp.media = dmedia.p;
    kp.gprop = dprop.as<float>();
    kp.field = dfield.as<float>();
    kp.field2 = dfield2.as<float>();
    kp.fieldk = dfieldk.as<int32_t>();
    kp.field64 = dfield64.as<double>();
    kp.fieldlen = fieldlen;
    kp.genergy = dgenergy.as<float>();
    kp.seeds = dseeds.as<uint32_t>();
    kp.detdata = ddet.as<float>();
    kp.detcount = dcount.as<uint32_t>();
    kp.counters = g_counters ? dcounters.as<uint64_t>() : nullptr;
    kp.srcpattern = dpattern.as<float>();
    kp.issaveref = cfg->issaveref ? 1u : 0u;
    kp.saveseed = saveseed ? 1u : 0u;
    kp.seeddata = dsaved.as<uint32_t>();
    kp.replay = replay ? 1 : 0;
    kp.replaydet = cfg->replaydet;
    kp.replayseed = drseed.as<uint32_t>();
    kp.replayweight = drweight.as<float>();
    kp.replaytof = drtof.as<float>();
    kp.replaydetid = drdetid.as<int32_t>();
    kp.replayindex = g_replayindex;
    kp.nmaps_mask = nrep - 1;
    kp.refill = refill > 1 ? 1u : 0u;
    kp.nlanes = nlanes;
    kp.nvthreads = (uint32_t)std::min<size_t>(nthread, (size_t)threadphoton * nthread + oddphotons);
    kp.omega = cfg->omega;
    kp.rfshift = rf ? (int32_t)(maxgate * blocks) : 0;

    /* per lane in LDS: the partial paths, then (seed saving) the launch RNG state */
    const size_t lane_lds = (partialdata + (saveseed ? 4 : 0)) * sizeof(float);
    unsigned int block = (unsigned int)G.autoblock;
    /* autopilot with the hot-cell box: one workgroup per CU holding all of
     * the CU's waves, so the box takes the CU's whole LDS instead of an
     * equal share of it between four workgroups */
    /* blocked tally: per lane 9 words of block and one queue word */
    const size_t lane_blk = tblock ? 40 : 0;
    if (cfg->autopilot && g_hotbox && tally != MCX_TALLY_F32_SHADOW && maxgate > 0) {
        const unsigned int whole = 64u * 4u * (unsigned int)waves_per_simd(numregs);
        if (whole > block && whole <= (unsigned int)maxthreads && nlanes % whole == 0 &&
            (size_t)ntable * 16 + (size_t)whole * lane_lds <= 16384 &&
            (size_t)ntable * 16 + (size_t)whole * (lane_lds + lane_blk) <= 65536)
            block = whole;
    }
    const unsigned int grid = nlanes / block;
    const size_t shmem0 = (size_t)ntable * 16 + (size_t)block * lane_lds;
    if (shmem0 > 65536) MCX_THROW(-1, "per-workgroup LDS request exceeds 64 KB; reduce nblocksize or the media count");
    size_t shmemb = shmem0;
    if (tblock) {
        const size_t blk_off = (shmem0 + 15) & ~(size_t)15;
        const size_t bq_off = (blk_off + (size_t)block * 36 + 15) & ~(size_t)15;
        if (bq_off + (size_t)block * 4 <= 65536) {
            kp.tblock = 1;
            kp.bnx = bnx;
            kp.bny = bny;
            kp.bnz = bnz;
            kp.blk_off = (uint32_t)blk_off;
            kp.bq_off = (uint32_t)bq_off;
            div_magic(kp.dimxy, &kp.hot_mxy, &kp.hot_shxy);
            div_magic(kp.nx, &kp.hot_mx, &kp.hot_shx);
            shmemb = bq_off + (size_t)block * 4;
        } else {
            tblock = false;
        }
    }
    kp.mapstride = tblock ? (uint64_t)(fieldlen / nvox) * bnz * bny * bnx * 8 : (uint64_t)fieldlen;
    /* replicas: the box's cells would mix the runs of a workgroup's lanes
     * (and each run's source cells see a K-th of the deposits) */
    const size_t shmem = nrep > 1 ? shmemb : set_hot_box(cfg, tally, block, numregs, shmemb, kp);

    if (cfg->flog && cfg->printnum >= 0) {
        fprintf(cfg->flog, "GPU=%d (%s) threadph=%d extra=%d np=%.0f nthread=%u maxgate=%u repetition=%d\n", gpuid + 1, G.name,
                threadphoton, oddphotons, (double)gpuphoton, nthread, maxgate, abs(cfg->respin));
        fflush(cfg->flog);
    }

    GlibcRand rnd((unsigned int)(cfg->seed > 0 ? cfg->seed + threadid : (int)time(0) + threadid));
    const size_t nseedw = (size_t)nthread * 4;
    uint32_t* seeds = pinned_seed_buffer(nseedw);
    std::vector<float> field(cfg->issave2pt ? fieldlen : 0, 0.f);
    std::vector<float> fieldacc((cfg->issave2pt && abs(cfg->respin) > 1) ? fieldlen : 0, 0.f);
    std::vector<float> pdet;
    double kernel_ms = 0.0, wall_ms = 0.0;
    int nlaunch = 0;
    uint32_t counts[4] = {0, 0, 0, 0};
    uint32_t errlanes = 0;
    EventPair ev, evc;
    /* launch statistics returned by mcx_amd_last_stats */
    auto record_stats = [&]() {
        uint64_t ctr[7] = {0, 0, 0, 0, 0, 0, 0};
        if (g_counters) HIP_CHECK(hipMemcpy(ctr, dcounters.p, sizeof(ctr), hipMemcpyDeviceToHost));
        std::lock_guard<std::mutex> lk(g_statmu);
        const double v[18] = {kernel_ms, wall_ms, (double)ctr[0], (double)ctr[1], (double)ctr[2], (double)ctr[3],
                              (double)numregs, (double)nthread, (double)block, (double)nlaunch, (double)errlanes,
                              (double)kp.hot_cells, (double)shmem, (double)ctr[4], tblock ? 1.0 : 0.0,
                              (double)ctr[5], (double)ctr[6], (double)nlanes};
        std::copy(v, v + 18, g_stats);
    };

    for (unsigned int timegate = 0; timegate < totalgates; timegate += maxgate) {
        kp.twin0 = cfg->tstart + cfg->tstep * timegate;
        kp.twin1 = kp.twin0 + cfg->tstep * maxgate;
        for (int iter = 0; iter < abs(cfg->respin); iter++) {
            if (g_rep) HIP_CHECK(hipEventRecord(evc.a, st.s));
            if (tally == MCX_TALLY_F64 || tally == MCX_TALLY_F64_XCD) {
                HIP_CHECK(hipMemsetAsync(dfield64.p, 0, dfield64.bytes, st.s));
            } else {
                HIP_CHECK(hipMemsetAsync(dfield.p, 0, dfield.bytes, st.s));
                HIP_CHECK(hipMemsetAsync(dfield2.p, 0, dfield2.bytes, st.s));
                if (dfieldk.p) HIP_CHECK(hipMemsetAsync(dfieldk.p, 0, dfieldk.bytes, st.s));
            }
            if (savedet) HIP_CHECK(hipMemsetAsync(ddet.p, 0, ddet.bytes, st.s));
            HIP_CHECK(hipMemsetAsync(dcount.p, 0, dcount.bytes, st.s));
            rnd.seed_words(seeds, nseedw);
            HIP_CHECK(hipMemcpyAsync(dseeds.p, seeds, nseedw * 4, hipMemcpyHostToDevice, st.s));
            HIP_CHECK(hipMemcpyAsync(dkp.p, &kp, sizeof(kp), hipMemcpyHostToDevice, st.s));
            HIP_CHECK(hipStreamSynchronize(st.s));

            auto tic = std::chrono::steady_clock::now();
            HIP_CHECK(hipEventRecord(ev.a, st.s));
            HIP_CHECK(mcxk_launch_transport(dkp.as<KParam>(), ispencil, isref > 0, tally, media32, grid, block, shmem, st.s));
            HIP_CHECK(hipEventRecord(ev.b, st.s));
            HIP_CHECK(hipMemcpyAsync(counts, dcount.p, sizeof(counts), hipMemcpyDeviceToHost, st.s));
            HIP_CHECK(hipStreamSynchronize(st.s));
            auto toc = std::chrono::steady_clock::now();
            float ms = 0.f;
            HIP_CHECK(hipEventElapsedTime(&ms, ev.a, ev.b));
            kernel_ms += ms;
            wall_ms += std::chrono::duration<double, std::milli>(toc - tic).count();
            nlaunch++;
            errlanes += counts[1];
            uint32_t detected = counts[0];
            if (g_rep) {
                finish_replicas(cfg, *g_rep, dgenergy, dfield64, nthread, nvox, (uint32_t)(fieldlen / nvox),
                                tblock ? bnx : 0u, bny, bnz, ev, evc, wall_ms, st.s);
                if (counts[1]) MCX_THROW(-2, "transport kernel: a lane stopped on an internal error");
                cfg->runtime = (unsigned int)llround(wall_ms);
                record_stats();
                return;
            }

            if (savedet) {
                if (detected > cfg->maxdetphoton) {
                    if (cfg->flog)
                        fprintf(cfg->flog, "WARNING: the detected photon number is more than what your have specified (%u > %u), please use the -H option to specify a greater number\n",
                                detected, cfg->maxdetphoton);
                } else if (cfg->flog) {
                    fprintf(cfg->flog, "detected %u photons, total: %.0f\n", detected, (double)cfg->detectedcount + detected);
                }
                cfg->his.detected += detected;
                detected = std::min(detected, cfg->maxdetphoton);
                if (cfg->exportdetected) {
                    cfg->exportdetected = (float*)realloc(cfg->exportdetected,
                                                          std::max<size_t>(1, (cfg->detectedcount + detected) * (size_t)hostdetreclen) * sizeof(float));
                    if (detected)
                        HIP_CHECK(hipMemcpyAsync(cfg->exportdetected + cfg->detectedcount * hostdetreclen, ddet.p,
                                                 (size_t)detected * hostdetreclen * sizeof(float), hipMemcpyDeviceToHost, st.s));
                    if (saveseed) {
                        cfg->seeddata = realloc(cfg->seeddata, std::max<size_t>(1, (cfg->detectedcount + detected) * (size_t)16));
                        if (!cfg->seeddata) MCX_THROW(-1, "out of host memory for photon seeds");
                        if (detected)
                            HIP_CHECK(hipMemcpyAsync((char*)cfg->seeddata + cfg->detectedcount * 16, dsaved.p, (size_t)detected * 16,
                                                     hipMemcpyDeviceToHost, st.s));
                    }
                    cfg->detectedcount += detected;
                }
            }
            if (cfg->issave2pt) {
                const void* a = (tally == MCX_TALLY_F64 || tally == MCX_TALLY_F64_XCD) ? dfield64.p : dfield.p;
                if (tblock)
                    HIP_CHECK(mcxk_launch_merge_blocked(dmerged.as<float>(), dfield64.as<double>(), cfg->dim.x, cfg->dim.y,
                                                        cfg->dim.z, bnx, bny, bnz, fieldlen, st.s));
                else
                    HIP_CHECK(mcxk_launch_merge(tally, dmerged.as<float>(), a, dfield2.p, dfieldk.p, fieldlen, st.s));
                HIP_CHECK(hipMemcpyAsync(field.data(), dmerged.p, fieldlen * sizeof(float), hipMemcpyDeviceToHost, st.s));
                HIP_CHECK(hipStreamSynchronize(st.s));
                if (abs(cfg->respin) > 1) {
#pragma omp parallel for schedule(static)
                    for (long long i = 0; i < (long long)fieldlen; i++) fieldacc[i] += field[i];
                }
            }
            HIP_CHECK(hipStreamSynchronize(st.s));
        }
        if (abs(cfg->respin) > 1 && cfg->issave2pt) field = fieldacc;

        std::vector<float> energy((size_t)nthread * 2);
        HIP_CHECK(hipMemcpyAsync(energy.data(), dgenergy.p, energy.size() * sizeof(float), hipMemcpyDeviceToHost, st.s));
        HIP_CHECK(hipStreamSynchronize(st.s));
        for (unsigned int i = 0; i < nthread; i++) {
            cfg->energyesc += energy[2 * i];
            cfg->energytot += energy[2 * i + 1];
        }
        if (cfg->exportfield && cfg->issave2pt) {
#pragma omp parallel for schedule(static)
            for (long long i = 0; i < (long long)fieldlen; i++) cfg->exportfield[i] += field[i];
        }
        if (kp.twin1 < cfg->tend) HIP_CHECK(hipMemsetAsync(dgenergy.p, 0, dgenergy.bytes, st.s));
        if (abs(cfg->respin) > 1) std::fill(fieldacc.begin(), fieldacc.end(), 0.f);
    }
    cfg->runtime = (unsigned int)llround(wall_ms);

    record_stats();

    /* normalisation (src/mcx_core.cu:3823-3925) */
    if (cfg->issave2pt && cfg->isnormalized) {
        float scale = 1.f;
        const float Vvox = cfg->steps.x * cfg->steps.y * cfg->steps.z;
        cfg->energyabs += cfg->energytot - cfg->energyesc;
        if (cfg->outputtype == otFlux || cfg->outputtype == otFluence) {
            scale = (float)(cfg->unitinmm / (cfg->energytot * Vvox * cfg->tstep));
            if (cfg->outputtype == otFluence) scale *= cfg->tstep;
        } else if (cfg->outputtype == otEnergy || cfg->outputtype == otL) {
            scale = (float)(1.f / cfg->energytot);
        }
        bool done = false;
        if (replay) {
            /* src/mcx_core.cu:3842-3889: replay tallies are divided by the
             * summed detected weight (per detector with replaydet -1); path
             * tallies (Jacobian, WL x TOF) are in grid units and also scaled
             * by unitinmm */
            const bool pathtype = cfg->outputtype == otJacobian || cfg->outputtype == otRF || cfg->outputtype == otWLTOF;
            if (cfg->replaydet == -1) {
                for (unsigned int det = 1; det <= cfg->detnum; det++) {
                    if (cfg->isnormalized != 2) {
                        scale = 0.f;
                        for (size_t i = 0; i < gpuphoton; i++)
                            if ((unsigned int)(cfg->replay.detid[i] & 0xFFFF) == det) scale += cfg->replay.weight[i];
                        if (scale > 0.f) {
                            scale = 1.0f / scale;
                            if (pathtype) scale = cfg->unitinmm * scale;
                        }
                    } else if (pathtype) {
                        scale = cfg->unitinmm;
                    }
                    if (cfg->flog) fprintf(cfg->flog, "normalization factor for detector %u alpha=%f\n", det, scale);
                    mcx_normalize(cfg->exportfield + (size_t)(det - 1) * nvox * maxgate, scale, nvox * maxgate, cfg->isnormalized, 0, 1);
                    if (rf)
                        mcx_normalize(cfg->exportfield + planelen + (size_t)(det - 1) * nvox * maxgate, scale, nvox * maxgate,
                                      cfg->isnormalized, 0, 1);
                }
                done = true;
            } else {
                scale = 0.f;
                for (size_t i = 0; i < gpuphoton; i++) scale += cfg->replay.weight[i];
                if (scale > 0.f) scale = cfg->unitinmm / scale;
            }
        }
        cfg->normalizer = scale;
        cfg->his.normalizer = scale;
        if (!done) {
            if (cfg->flog) fprintf(cfg->flog, "source 1, normalization factor alpha=%f\n", scale);
            mcx_normalize(cfg->exportfield, scale, fieldlen, cfg->isnormalized, 0, 1);
        }
    }
    if (cfg->parentid == mpStandalone) {
        if (cfg->issave2pt) mcx_savedata(cfg->exportfield, planelen, cfg);
        if (cfg->issavedet && cfg->exportdetected) {
            cfg->his.unitinmm = cfg->unitinmm;
            cfg->his.savedphoton = (unsigned int)cfg->detectedcount;
            cfg->his.totalphoton = (unsigned int)cfg->nphoton;
            if (cfg->issaveseed) cfg->his.seedbyte = 16;
            cfg->his.detected = (unsigned int)cfg->detectedcount;
            mcx_savedetphoton(cfg->exportdetected, cfg->seeddata, (int)cfg->detectedcount, saveseed ? 16 : 0, cfg);
        }
    }
    if (cfg->flog) {
        const double np = (double)cfg->nphoton * ((cfg->respin > 1) ? cfg->respin : 1);
        fprintf(cfg->flog, "simulated %.0f photons (%.0f) with %u threads (repeat x%d)\nMCX simulation speed: %.2f photon/ms\n", np, np,
                nthread, abs(cfg->respin), np / std::max(1e-3, wall_ms));
        fprintf(cfg->flog, "total simulated energy: %.2f\tabsorbed: %5.5f%%\n(loss due to initial specular reflection is excluded in the total)\n",
                cfg->energytot, (cfg->energytot - cfg->energyesc) / cfg->energytot * 100.f);
        fflush(cfg->flog);
    }
    cfg->energyabs = cfg->energytot - cfg->energyesc;
    if (errlanes)
        MCX_THROW(-2, "transport kernel: a lane stopped on an internal error (a photon outside the volume without a relaunch, "
                      "or a history beyond the step limit)");
}

static int list_gpu_impl(Config* cfg, GPUInfo** info) {
    int count = 0;
    hipError_t e = hipGetDeviceCount(&count);
    if (e != hipSuccess || count == 0) {
        if (cfg->flog) fprintf(cfg->flog, "ERROR: No HIP-capable GPU device found\n");
        *info = NULL;
        return 0;
    }
    *info = (GPUInfo*)calloc((size_t)count, sizeof(GPUInfo));
    if (cfg->gpuid && cfg->gpuid > count) {
        if (cfg->flog) fprintf(cfg->flog, "ERROR: Specified GPU ID is out of range\n");
        free(*info);
        *info = NULL;
        return 0;
    }
    int activedev = 0;
    for (int dev = 0; dev < count; dev++) {
        hipDeviceProp_t dp;
        HIP_CHECK(hipGetDeviceProperties(&dp, dev));
        if (cfg->isgpuinfo == 3) {
            activedev++;
        } else if (cfg->deviceid[dev] == '1') {
            cfg->deviceid[dev] = '\0';
            cfg->deviceid[activedev] = (char)(dev + 1);
            activedev++;
        }
        GPUInfo& g = (*info)[dev];
        strncpy(g.name, dp.name, MCX_MAX_SESSION_LENGTH - 1);
        g.id = dev + 1;
        g.devcount = count;
        g.major = dp.major;
        g.minor = dp.minor;
        g.globalmem = dp.totalGlobalMem;
        g.constmem = dp.totalConstMem;
        g.sharedmem = dp.sharedMemPerBlock;
        g.regcount = dp.regsPerBlock;
        g.clock = dp.clockRate;
        g.sm = dp.multiProcessorCount;
        g.core = dp.multiProcessorCount * 4 * 16; /* 4 SIMDs x 16 lanes per CU */
        g.maxmpthread = dp.maxThreadsPerMultiProcessor;
        g.maxgate = cfg->maxgate;
        g.autoblock = 256;
        g.autothread = g.sm * 4 * 8 * 64;
        if (cfg->isgpuinfo && cfg->flog) {
            fprintf(cfg->flog, "=============================   GPU Information  ================================\n");
            fprintf(cfg->flog, "Device %d of %d:\t\t%s (%s)\n", g.id, g.devcount, g.name, dp.gcnArchName);
            fprintf(cfg->flog, "Global Memory:\t\t%.0f B\nCompute Units:\t\t%d\nClock Speed:\t\t%.2f GHz\n", (double)g.globalmem, g.sm,
                    g.clock * 1e-6f);
            fprintf(cfg->flog, "Auto-thread:\t\t%d\nAuto-block:\t\t%d\n", g.autothread, g.autoblock);
        }
    }
    if (cfg->isgpuinfo == 2 && cfg->parentid == mpStandalone) exit(0);
    if (activedev < MCX_MAX_DEVICE) cfg->deviceid[activedev] = '\0';
    return activedev;
}

}  // namespace mcx

using namespace mcx;

extern "C" {

void mcx_run_simulation(Config* cfg, GPUInfo* gpu) {
    EntryGuard g;
    GPUInfo* own = NULL;
    try {
        if (gpu == NULL) {
            if (list_gpu_impl(cfg, &own) == 0) MCX_THROW(-1, "no active GPU device");
            gpu = own;
        }
        run_impl(cfg, gpu);
    } catch (const Error& e) {
        if (own) free(own);
        handle_error(e, cfg);
        return;
    }
    if (own) free(own);
}

int mcx_list_gpu(Config* cfg, GPUInfo** info) {
    EntryGuard g;
    try {
        return list_gpu_impl(cfg, info);
    } catch (const Error& e) {
        handle_error(e, cfg);
        return 0;
    }
}

int mcx_rng_test(unsigned int seed, float* out, unsigned int n) {
    EntryGuard g;
    try {
        GlibcRand r(seed);
        uint32_t seeds[4];
        r.seed_words(seeds, 4);
        Stream st;
        DevMem dseed, dout;
        dseed.alloc(sizeof(seeds));
        dout.alloc((size_t)n * sizeof(float));
        HIP_CHECK(hipMemcpyAsync(dseed.p, seeds, sizeof(seeds), hipMemcpyHostToDevice, st.s));
        HIP_CHECK(mcxk_launch_rngtest(dout.as<float>(), dseed.as<uint32_t>(), n, st.s));
        HIP_CHECK(hipMemcpyAsync(out, dout.p, (size_t)n * sizeof(float), hipMemcpyDeviceToHost, st.s));
        HIP_CHECK(hipStreamSynchronize(st.s));
        return 0;
    } catch (const Error& e) {
        record_error(e.id, e.msg.c_str());
        return e.id ? e.id : -1;
    }
}

int mcx_amd_set_tally_mode(int mode) {
    if (mode < 0 || mode > 3) return -1;
    int old = g_tally_mode;
    g_tally_mode = mode;
    return old;
}

void mcx_amd_set_counters(int enable) { g_counters = enable ? 1 : 0; }

int mcx_amd_set_replay_index(int reference) {
    int old = g_replayindex;
    g_replayindex = reference ? 1 : 0;
    return old;
}

int mcx_amd_seed_words(unsigned int seed, unsigned int* out, const unsigned long long* counts, int ncalls) {
    EntryGuard g;
    try {
        if (!out || !counts || ncalls < 0) MCX_THROW(-1, "seed words: NULL output or negative call count");
        GlibcRand r(seed);
        for (int i = 0; i < ncalls; i++) {
            r.seed_words(out, (size_t)counts[i]);
            out += counts[i];
        }
        return 0;
    } catch (const Error& e) {
        record_error(e.id, e.msg.c_str());
        return e.id;
    }
}

int mcx_amd_run_replicas(Config* cfg, GPUInfo* gpu, int k, float* cw, double* energy, double* stats) {
    EntryGuard g;
    GPUInfo* own = NULL;
    ReplicaReq req{k, cw, energy, stats};
    try {
        if (k < 1 || k > 64 || (k & (k - 1))) MCX_THROW(-1, "the replica count must be a power of two from 1 to 64");
        if (!cw || !energy || !stats) MCX_THROW(-1, "replica outputs can not be NULL");
        if (gpu == NULL) {
            if (list_gpu_impl(cfg, &own) == 0) MCX_THROW(-1, "no active GPU device");
            gpu = own;
        }
        g_rep = &req;
        run_impl(cfg, gpu);
        g_rep = nullptr;
    } catch (const Error& e) {
        g_rep = nullptr;
        if (own) free(own);
        record_error(e.id, e.msg.c_str());
        return e.id ? e.id : -1;
    }
    if (own) free(own);
    return 0;
}

int mcx_amd_set_refill(int mode) {
    if (mode < 0 || mode > 2) return -1;
    int old = g_refill;
    g_refill = mode;
    return old;
}

int mcx_amd_set_tally_blocks(int mode) {
    if (mode < 0 || mode > 2) return -1;
    int old = g_tblock;
    g_tblock = mode;
    return old;
}

int mcx_amd_set_hot_box(int enable) {
    int old = g_hotbox;
    g_hotbox = enable ? 1 : 0;
    return old;
}

int mcx_amd_last_stats(double* out, int n) {
    std::lock_guard<std::mutex> lk(g_statmu);
    int k = std::min(n, 18);
    for (int i = 0; i < k; i++) out[i] = g_stats[i];
    return k;
}

}  // extern "C"
