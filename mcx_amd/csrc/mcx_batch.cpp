/*
 * mcx_batch.cpp -- many independent simulations in one persistent launch.
 *
 * The reference runs one simulation per mcx_run_simulation call, one kernel
 * launch per gate group and respin (src/mcx_core.cu:3344-3392).  Here every
 * instance's launch constants go into a device array of KParam blocks, its
 * virtual threads are cut into chunks of 64 (one wave), and a persistent
 * grid sized to the device's resident wave count pulls chunks from a global
 * counter, most expensive instance first.  Each virtual thread is seeded and
 * runs exactly as in the single-simulation path with the same thread count,
 * so a batched instance reproduces that run lane for lane.
 *
 * One execution is: zero the tally/energy arena (fill kernel, no memset, so
 * the sequence can be captured), transport, per-instance FP64 energy sums,
 * FP64 merge of the tally halves with normalisation.  With use_graph the
 * sequence is captured once into a hipGraph and replayed.
 *
 * The launch shape (virtual threads per instance via photons-per-lane, and
 * the workgroup size) is chosen by a measured autotuner when not given.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#include <unistd.h>

#include <algorithm>
#include <chrono>
#include <climits>
#include <map>
#include <mutex>
#include <numeric>
#include <vector>

#include "../../include/mcx_amd_ext.h"
#include "mcx_device.h"

extern "C" hipError_t mcxk_launch_batch(int tally, const KParam* params, const uint32_t* chunk_start, const uint32_t* order,
                                        uint32_t ninst, uint32_t total_chunks, uint32_t maxtab, uint32_t boxcells,
                                        uint32_t* queue, unsigned int grid, unsigned int block, hipStream_t stream);
extern "C" hipError_t mcxk_batch_attrs(int tally, int* numregs);
extern "C" hipError_t mcxk_launch_batch_post(int tally, const KParam* params, BatchInst* bi, uint32_t ninst, hipStream_t stream);
extern "C" hipError_t mcxk_fill_zero16(void* p, size_t bytes, hipStream_t stream);
extern "C" hipError_t mcxk_launch_batch_cw(const KParam* params, const BatchInst* bi, uint32_t ninst, hipStream_t stream);

namespace mcx {

static size_t align256(size_t n) { return (n + 255) & ~(size_t)255; }

/* With MCX_AMD_VERBOSE set, each phase of a batched call is reported on
 * stderr with the seconds since the call started, so a long call shows
 * where its time goes. */
static void phase(const char* what, const std::chrono::steady_clock::time_point& t0) {
    static const bool on = getenv("MCX_AMD_VERBOSE") != nullptr;
    if (!on) return;
    fprintf(stderr, "[mcx_run_batch] %s at %.2f s\n", what,
            std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count());
    fflush(stderr);
}

struct BatchShape {
    int photons_per_lane = 0;  /* 0: each Config's own nthread (autopilot off) */
    int block = 256;
    int boxcells = 0;          /* hot-cell box cells per wave (tally cache in LDS); 0: none */
};

/* Largest hot-cell box per wave: the LDS a wave can take at the batched
 * kernel's full occupancy (and within 64 KB per workgroup), less its
 * property table, at 8 bytes per cell. */
static int batch_box_budget(int device, int vgprs, unsigned int block, uint32_t maxtab) {
    const unsigned int waves_cu = 4u * (unsigned int)waves_per_simd(vgprs);
    const size_t per_wave = lds_share(device, waves_cu, 65536 / std::max(1u, block / 64u));
    const size_t tab = (size_t)maxtab * 16;
    return per_wave > tab ? (int)((per_wave - tab) / 8) : 0;
}

/* Host-side description of one instance */
struct InstPlan {
    Config* cfg;
    size_t nvox, fieldlen;
    unsigned int gates, nthread;
    int threadphoton, oddphotons;
    double cost;
    size_t off_media, off_prop, off_seed, off_tally, off_tally2, off_tallyk, off_energy, off_genergy, off_out;
};

class BatchRun {
   public:
    BatchRun(Config** cfgs, int n, int tally, const BatchShape& shape, double photon_scale, int device)
        : cfgs_(cfgs), n_(n), tally_(tally), shape_(shape), scale_(photon_scale), device_(device) {}
    ~BatchRun() {
        if (gexec_) (void)hipGraphExecDestroy(gexec_);
        if (graph_) (void)hipGraphDestroy(graph_);
    }

    void setup();
    /* a new workgroup size and tally cache for the next runs, on the same
     * arenas (the autotuner's trials) */
    void reshape(int block, int boxcells);
    void enqueue(hipStream_t s);
    float run(hipStream_t s, bool graph, int repeat, float* kernel_ms);
    /* energies always; fields 1: every gate's fluence, 2: the
     * time-integrated map (nvox floats), 0: none.  Returns the device time
     * of the time-integrated reduction in ms (0 without it). */
    float download(hipStream_t s, int fields);
    void read_counters(hipStream_t s, double* out, double* wave_iterations);

    double photons = 0.0;
    unsigned long long lanes = 0;
    int vgprs = 0;
    int box_cells() const { return (int)boxcells_; }

   private:
    Config** cfgs_;
    int n_, tally_;
    BatchShape shape_;
    double scale_;
    int device_;
    std::vector<InstPlan> plan_;
    std::vector<uint32_t> order_, chunk_start_;
    std::vector<KParam> kps_;
    std::vector<BatchInst> bis_;
    uint32_t total_chunks_ = 0, maxtab_ = 0, boxcells_ = 0;
    unsigned int grid_ = 0;
    DevMem dmedia_, dprop_, dseed_, dzero_, dout_, dkp_, dchunk_, dorder_, dbi_, dcw_;
    size_t queue_off_ = 0, counters_off_ = 0;
    hipGraph_t graph_ = nullptr;
    hipGraphExec_t gexec_ = nullptr;
};

static void check_batch_config(const Config* cfg, int i) {
    char msg[256];
    check_supported(cfg);
    if (cfg->issavedet) {
        snprintf(msg, sizeof(msg), "batch instance %d: detected-photon output is not supported in a batched launch", i);
        MCX_THROW(-1, msg);
    }
    if (cfg->issaveref) {
        snprintf(msg, sizeof(msg), "batch instance %d: diffuse reflectance output is not batched", i);
        MCX_THROW(-1, msg);
    }
    if (cfg->mediabyte > 4) {
        snprintf(msg, sizeof(msg), "batch instance %d: a batched launch takes label volumes only", i);
        MCX_THROW(-1, msg);
    }
    if (cfg->srctype == MCX_SRC_PATTERN || cfg->srctype == MCX_SRC_PATTERN3D) {
        snprintf(msg, sizeof(msg), "batch instance %d: pattern sources are not batched", i);
        MCX_THROW(-1, msg);
    }
    if (cfg->seed == MCX_SEED_FROM_FILE || cfg->issaveseed) {
        snprintf(msg, sizeof(msg), "batch instance %d: photon replay and seed saving are not batched", i);
        MCX_THROW(-1, msg);
    }
    if (cfg->respin != 1) {
        snprintf(msg, sizeof(msg), "batch instance %d: respin must be 1 in a batched launch", i);
        MCX_THROW(-1, msg);
    }
    if (!cfg->issave2pt) {
        snprintf(msg, sizeof(msg), "batch instance %d: a batched launch always produces the fluence (issave2pt)", i);
        MCX_THROW(-1, msg);
    }
    if (cfg->debuglevel & MCX_DEBUG_RNG) MCX_THROW(-1, "the RNG test mode cannot be batched");
    if (cfg->nphoton == 0) {
        snprintf(msg, sizeof(msg), "batch instance %d: no photons", i);
        MCX_THROW(-1, msg);
    }
    if (cfg->medianum + cfg->detnum > MCX_MAX_PROP_AND_DETECTORS) MCX_THROW(-1, "too many media in a batch instance");
}

void BatchRun::setup() {
    HIP_CHECK(hipSetDevice(device_));
    plan_.resize(n_);
    size_t media_b = 0, prop_b = 0, seed_b = 0, zero_b = 0, out_b = 0;
    for (int i = 0; i < n_; i++) {
        Config* cfg = cfgs_[i];
        check_batch_config(cfg, i);
        InstPlan& P = plan_[i];
        P.cfg = cfg;
        P.nvox = (size_t)cfg->dim.x * cfg->dim.y * cfg->dim.z;
        P.gates = (unsigned int)((cfg->tend - cfg->tstart) / cfg->tstep + 0.5);
        if (P.gates == 0) P.gates = 1;
        P.fieldlen = P.nvox * P.gates;
        for (size_t k = 0; k < P.nvox; k++)
            if ((cfg->vol[k] & MCX_MED_MASK) > 127 || (cfg->vol[k] & MCX_DET_MASK))
                MCX_THROW(-1, "batched launch: labels must be below 128 and carry no detector mask");
        size_t nph = (size_t)std::max(1.0, floor((double)cfg->nphoton * scale_));
        if (shape_.photons_per_lane > 0) {
            size_t nt = (nph + shape_.photons_per_lane - 1) / shape_.photons_per_lane;
            nt = std::max<size_t>(64, (nt + 63) / 64 * 64);
            P.nthread = (unsigned int)nt;
        } else {
            if (cfg->nthread <= 0) MCX_THROW(-1, "batched launch: set nthread or photons_per_lane");
            P.nthread = (unsigned int)cfg->nthread;
        }
        P.threadphoton = (int)(nph / P.nthread);
        P.oddphotons = (int)(nph - (size_t)P.threadphoton * P.nthread);
        photons += (double)nph;
        lanes += P.nthread;
        /* relative cost of one chunk: photons per lane x an estimate of the
         * steps per photon (scattering lengths across the volume, capped by
         * absorption), used only to order the work queue */
        double mus = 0.0, mua = 0.0;
        for (unsigned int m = 1; m < cfg->medianum; m++) {
            mus = std::max(mus, (double)cfg->prop[m].mus);
            mua = std::max(mua, (double)cfg->prop[m].mua);
        }
        const double edge = cbrt((double)P.nvox);
        const double steps = std::min(edge * edge * mus * mus + edge, mus / std::max(mua, 1e-6) + edge);
        P.cost = ((double)P.threadphoton + (P.oddphotons > 0)) * (1.0 + steps);
        P.off_media = media_b;
        media_b += align256(P.nvox);
        P.off_prop = prop_b;
        prop_b += align256((size_t)16 * (cfg->medianum + cfg->detnum));
        P.off_seed = seed_b;
        seed_b += align256((size_t)16 * P.nthread);
        P.off_tally = zero_b;
        zero_b += align256(P.fieldlen * (tally_ == MCX_TALLY_F64 ? 8 : 4));
        P.off_tally2 = zero_b;
        if (tally_ != MCX_TALLY_F64) zero_b += align256(P.fieldlen * 4);
        P.off_tallyk = zero_b;
        if (tally_ == MCX_TALLY_F32_HILO) zero_b += align256(P.fieldlen * 4);
        P.off_genergy = zero_b;
        zero_b += align256((size_t)8 * P.nthread);
        P.off_energy = zero_b;
        zero_b += 256;
        P.off_out = out_b;
        out_b += align256(P.fieldlen * 4);
        maxtab_ = std::max(maxtab_, cfg->medianum + cfg->detnum);
    }
    queue_off_ = zero_b;
    zero_b += 256;
    counters_off_ = zero_b;
    zero_b += 256;

    /* work queue: instances by decreasing chunk cost */
    order_.resize(n_);
    std::iota(order_.begin(), order_.end(), 0u);
    std::stable_sort(order_.begin(), order_.end(), [&](uint32_t a, uint32_t b) { return plan_[a].cost > plan_[b].cost; });
    chunk_start_.resize(n_);
    total_chunks_ = 0;
    for (int k = 0; k < n_; k++) {
        chunk_start_[k] = total_chunks_;
        total_chunks_ += (plan_[order_[k]].nthread + 63) / 64;
    }

    HIP_CHECK(mcxk_batch_attrs(tally_, &vgprs));

    /* host staging */
    Stream st;
    std::vector<unsigned char> hmedia(media_b, 0);
    std::vector<float> hprop(prop_b / 4, 0.f);
    std::vector<uint32_t> hseed(seed_b / 4, 0u);
#pragma omp parallel for schedule(dynamic, 1)
    for (int i = 0; i < n_; i++) {
        const InstPlan& P = plan_[i];
        const Config* cfg = P.cfg;
        unsigned char* m = hmedia.data() + P.off_media;
        for (size_t k = 0; k < P.nvox; k++) m[k] = (unsigned char)(cfg->vol[k] & 0x7Fu);
        memcpy(hprop.data() + P.off_prop / 4, cfg->prop, 16 * cfg->medianum);
        if (cfg->detnum) memcpy(hprop.data() + P.off_prop / 4 + 4 * cfg->medianum, cfg->detpos, 16 * cfg->detnum);
        /* the single-simulation path's seeding: srand(seed + 0), then
         * Pseed[i] = (rand()<<16) | (rand()<<1) | (rand()>>14) */
        GlibcRand rnd((unsigned int)(cfg->seed > 0 ? cfg->seed : (int)time(0)));
        rnd.seed_words(hseed.data() + P.off_seed / 4, (size_t)P.nthread * 4);
    }
    dmedia_.alloc(media_b);
    dprop_.alloc(prop_b);
    dseed_.alloc(seed_b);
    dzero_.alloc(zero_b);
    dout_.alloc(out_b);
    dkp_.alloc(sizeof(KParam) * n_);
    dchunk_.alloc(4 * n_);
    dorder_.alloc(4 * n_);
    dbi_.alloc(sizeof(BatchInst) * n_);

    char* zb = dzero_.as<char>();
    kps_.assign(n_, KParam{});
    bis_.assign(n_, BatchInst{});
    std::vector<KParam>& kps = kps_;
    std::vector<BatchInst>& bis = bis_;
    for (int i = 0; i < n_; i++) {
        const InstPlan& P = plan_[i];
        const Config* cfg = P.cfg;
        unsigned int is2d = (cfg->dim.x == 1 ? 1 : (cfg->dim.y == 1 ? 2 : (cfg->dim.z == 1 ? 3 : 0)));
        if (is2d) is2d = is2d * ((cfg->dim.x > 1) + (cfg->dim.y > 1) + (cfg->dim.z > 1) == 2);
        KParam& kp = kps[i];
        fill_kparam(cfg, P.gates, is2d, kp);
        kp.twin0 = cfg->tstart;
        kp.twin1 = kp.twin0 + cfg->tstep * P.gates;
        kp.threadphoton = P.threadphoton;
        kp.oddphotons = P.oddphotons;
        kp.nthread = P.nthread;
        kp.media = dmedia_.as<char>() + P.off_media;
        kp.gprop = (const float*)(dprop_.as<char>() + P.off_prop);
        if (tally_ == MCX_TALLY_F64) {
            kp.field64 = (double*)(zb + P.off_tally);
        } else {
            kp.field = (float*)(zb + P.off_tally);
            kp.field2 = (float*)(zb + P.off_tally2);
            kp.fieldk = (int32_t*)(zb + P.off_tallyk);
        }
        kp.fieldlen = P.fieldlen;
        kp.genergy = (float*)(zb + P.off_genergy);
        kp.seeds = (const uint32_t*)(dseed_.as<char>() + P.off_seed);
        kp.detcount = (uint32_t*)(zb + P.off_energy + 64); /* error-exit counter ([1]) */
        kp.counters = counters_enabled() ? (uint64_t*)(zb + counters_off_) : nullptr;
        BatchInst& b = bis[i];
        memset(&b, 0, sizeof(b));
        b.out = (float*)(dout_.as<char>() + P.off_out);
        b.energy = (double*)(zb + P.off_energy);
        b.unitinmm = cfg->unitinmm;
        b.vvox = cfg->steps.x * cfg->steps.y * cfg->steps.z;
        b.tstep = cfg->tstep;
        b.outputtype = cfg->outputtype;
        b.normalize = cfg->isnormalized ? 1 : 0;
    }
    HIP_CHECK(hipMemcpyAsync(dmedia_.p, hmedia.data(), media_b, hipMemcpyHostToDevice, st.s));
    HIP_CHECK(hipMemcpyAsync(dprop_.p, hprop.data(), prop_b, hipMemcpyHostToDevice, st.s));
    HIP_CHECK(hipMemcpyAsync(dseed_.p, hseed.data(), seed_b, hipMemcpyHostToDevice, st.s));
    HIP_CHECK(hipMemcpyAsync(dchunk_.p, chunk_start_.data(), 4 * n_, hipMemcpyHostToDevice, st.s));
    HIP_CHECK(hipMemcpyAsync(dorder_.p, order_.data(), 4 * n_, hipMemcpyHostToDevice, st.s));
    HIP_CHECK(hipMemcpyAsync(dbi_.p, bis.data(), sizeof(BatchInst) * n_, hipMemcpyHostToDevice, st.s));
    HIP_CHECK(hipStreamSynchronize(st.s));
    reshape(shape_.block, shape_.boxcells);
}

void BatchRun::reshape(int block_size, int boxcells) {
    hipDeviceProp_t dp;
    HIP_CHECK(hipGetDeviceProperties(&dp, device_));
    const unsigned int block = (unsigned int)block_size;
    if (block < 64 || block > 1024 || block % 64) MCX_THROW(-1, "batched launch: block size must be a multiple of 64 in [64, 1024]");
    if ((size_t)(block / 64) * maxtab_ * 16 > 65536) MCX_THROW(-1, "batched launch: property tables exceed the LDS budget");
    if (boxcells < 0) MCX_THROW(-1, "batched launch: negative tally cache size");
    shape_.block = block_size;
    shape_.boxcells = boxcells;
    boxcells_ = (uint32_t)std::min(boxcells, batch_box_budget(device_, vgprs, block, maxtab_));
    if (boxcells_ < 64) boxcells_ = 0;
    const unsigned long long resident = (unsigned long long)dp.multiProcessorCount * 4 * waves_per_simd(vgprs) * 64;
    grid_ = (unsigned int)std::max<unsigned long long>(1, resident / block);
    /* each instance's hot-cell box lives in the wave's LDS slice; the
     * kernel passes the slice's offset */
    for (int i = 0; i < n_; i++) {
        hot_box_geometry(plan_[i].cfg, boxcells_, kps_[i]);
        kps_[i].hot_off = 0;
    }
    Stream st;
    HIP_CHECK(hipMemcpyAsync(dkp_.p, kps_.data(), sizeof(KParam) * n_, hipMemcpyHostToDevice, st.s));
    HIP_CHECK(hipStreamSynchronize(st.s));
    if (gexec_) {
        (void)hipGraphExecDestroy(gexec_);
        gexec_ = nullptr;
    }
    if (graph_) {
        (void)hipGraphDestroy(graph_);
        graph_ = nullptr;
    }
}

/* Physical memory currently available to this process, in bytes (0 when
 * unknown) */
static size_t host_bytes_available() {
    const long pages = sysconf(_SC_AVPHYS_PAGES), psz = sysconf(_SC_PAGESIZE);
    return (pages > 0 && psz > 0) ? (size_t)pages * (size_t)psz : 0;
}

void BatchRun::enqueue(hipStream_t s) {
    HIP_CHECK(mcxk_fill_zero16(dzero_.p, dzero_.bytes, s));
    HIP_CHECK(mcxk_launch_batch(tally_, dkp_.as<KParam>(), dchunk_.as<uint32_t>(), dorder_.as<uint32_t>(), (uint32_t)n_,
                                total_chunks_, maxtab_, boxcells_, (uint32_t*)(dzero_.as<char>() + queue_off_), grid_,
                                (unsigned int)shape_.block, s));
    HIP_CHECK(mcxk_launch_batch_post(tally_, dkp_.as<KParam>(), dbi_.as<BatchInst>(), (uint32_t)n_, s));
}

/* Executes the sequence `repeat` times; returns the last execution's time
 * in ms.  Without a graph, kernel_ms receives the transport kernel alone. */
float BatchRun::run(hipStream_t s, bool graph, int repeat, float* kernel_ms) {
    EventPair ev, kev;
    float total = 0.f;
    if (graph && !gexec_) {
        HIP_CHECK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
        try {
            enqueue(s);
        } catch (...) {
            hipGraph_t g = nullptr;
            (void)hipStreamEndCapture(s, &g);
            if (g) (void)hipGraphDestroy(g);
            throw;
        }
        HIP_CHECK(hipStreamEndCapture(s, &graph_));
        HIP_CHECK(hipGraphInstantiate(&gexec_, graph_, nullptr, nullptr, 0));
    }
    for (int r = 0; r < std::max(1, repeat); r++) {
        HIP_CHECK(hipEventRecord(ev.a, s));
        if (graph) {
            HIP_CHECK(hipGraphLaunch(gexec_, s));
        } else {
            HIP_CHECK(mcxk_fill_zero16(dzero_.p, dzero_.bytes, s));
            HIP_CHECK(hipEventRecord(kev.a, s));
            HIP_CHECK(mcxk_launch_batch(tally_, dkp_.as<KParam>(), dchunk_.as<uint32_t>(), dorder_.as<uint32_t>(), (uint32_t)n_,
                                        total_chunks_, maxtab_, boxcells_, (uint32_t*)(dzero_.as<char>() + queue_off_), grid_,
                                        (unsigned int)shape_.block, s));
            HIP_CHECK(hipEventRecord(kev.b, s));
            HIP_CHECK(mcxk_launch_batch_post(tally_, dkp_.as<KParam>(), dbi_.as<BatchInst>(), (uint32_t)n_, s));
        }
        HIP_CHECK(hipEventRecord(ev.b, s));
        HIP_CHECK(hipStreamSynchronize(s));
        HIP_CHECK(hipEventElapsedTime(&total, ev.a, ev.b));
        if (kernel_ms) {
            if (graph)
                *kernel_ms = total;
            else
                HIP_CHECK(hipEventElapsedTime(kernel_ms, kev.a, kev.b));
        }
    }
    return total;
}

float BatchRun::download(hipStream_t s, int fields) {
    float cw_ms = 0.f;
    if (fields == 2) {
        /* time-integrated maps reduced on the device: nvox floats per
         * instance instead of nvox x gates */
        size_t cw_b = 0;
        std::vector<size_t> off(n_);
        for (int i = 0; i < n_; i++) {
            off[i] = cw_b;
            cw_b += align256(plan_[i].nvox * 4);
        }
        if (!dcw_.p) dcw_.alloc(cw_b);
        for (int i = 0; i < n_; i++) bis_[i].cw = (float*)(dcw_.as<char>() + off[i]);
        HIP_CHECK(hipMemcpyAsync(dbi_.p, bis_.data(), sizeof(BatchInst) * n_, hipMemcpyHostToDevice, s));
        EventPair ev;
        HIP_CHECK(hipEventRecord(ev.a, s));
        HIP_CHECK(mcxk_launch_batch_cw(dkp_.as<KParam>(), dbi_.as<BatchInst>(), (uint32_t)n_, s));
        HIP_CHECK(hipEventRecord(ev.b, s));
        HIP_CHECK(hipEventSynchronize(ev.b));
        HIP_CHECK(hipEventElapsedTime(&cw_ms, ev.a, ev.b));
    }
    /* host buffers the library has to allocate: refuse before allocating
     * when they would not fit in the memory available */
    if (fields) {
        size_t need = 0;
        for (int i = 0; i < n_; i++)
            if (!plan_[i].cfg->exportfield) need += (fields == 2 ? plan_[i].nvox : plan_[i].fieldlen) * 4;
        const size_t avail = host_bytes_available();
        if (need && avail && need > avail / 10 * 8) {
            char msg[256];
            snprintf(msg, sizeof(msg), "batched launch: the fluence of %d instances needs %.1f GB of host memory, %.1f GB available",
                     n_, need / 1e9, avail / 1e9);
            MCX_THROW(-1, msg);
        }
    }
    std::vector<double> en((size_t)n_ * 32);
    std::vector<uint32_t> errs(n_);
    for (int i = 0; i < n_; i++) {
        const InstPlan& P = plan_[i];
        HIP_CHECK(hipMemcpyAsync(&en[(size_t)i * 32], dzero_.as<char>() + P.off_energy, 256, hipMemcpyDeviceToHost, s));
    }
    HIP_CHECK(hipStreamSynchronize(s));
    for (int i = 0; i < n_; i++) {
        const InstPlan& P = plan_[i];
        Config* cfg = P.cfg;
        const double* e = &en[(size_t)i * 32];
        uint32_t err;
        memcpy(&err, (const char*)e + 64 + 4, 4);
        if (err)
            MCX_THROW(-2, "batched transport kernel: a lane stopped on an internal error (a photon outside the volume without "
                          "a relaunch, or a history beyond the step limit)");
        cfg->energyesc = e[0];
        cfg->energytot = e[1];
        cfg->energyabs = e[1] - e[0];
        cfg->normalizer = (float)e[2];
        cfg->his.normalizer = (float)e[2];
        cfg->maxgate = P.gates;
        if (fields) {
            const size_t len = fields == 2 ? P.nvox : P.fieldlen;
            if (!cfg->exportfield) cfg->exportfield = (float*)calloc(len, sizeof(float));
            if (!cfg->exportfield) MCX_THROW(-1, "batched launch: out of host memory for the fluence");
            const char* src = fields == 2 ? (const char*)bis_[i].cw : dout_.as<char>() + P.off_out;
            HIP_CHECK(hipMemcpyAsync(cfg->exportfield, src, len * 4, hipMemcpyDeviceToHost, s));
        }
    }
    HIP_CHECK(hipStreamSynchronize(s));
    return cw_ms;
}

void BatchRun::read_counters(hipStream_t s, double* out, double* wave_iterations) {
    uint64_t c[7] = {0, 0, 0, 0, 0, 0, 0};
    if (counters_enabled()) {
        HIP_CHECK(hipMemcpyAsync(c, dzero_.as<char>() + counters_off_, sizeof(c), hipMemcpyDeviceToHost, s));
        HIP_CHECK(hipStreamSynchronize(s));
    }
    for (int i = 0; i < 4; i++) out[i] = (double)c[i];
    *wave_iterations = (double)c[6];
}

/* Autotuner.  Photons per lane: a wave runs as long as its slowest lane,
 * so more photons per lane balance the lanes of a chunk better, until the
 * chunks become too few to keep every resident wave busy to the end.  The
 * rule takes the largest candidate that leaves at least kChunksPerWave
 * chunks per resident wave at the batch's full photon counts (a reduced
 * copy of the batch cannot show this trade-off: it has fewer chunks, and
 * trials on one favour few photons per lane).  When no candidate leaves
 * that many the batch is small and bound by the longest chain of
 * histories one lane runs: {1, 2, 4, 8, 16} photons per lane are then
 * timed on the full batch (8 head-volume runs of 1.25e5 photons: 143 ms
 * at 16 photons per lane, 94 ms at 1; 8 of 5e5: 154 ms at 16, 227 ms at
 * 2, where the per-chunk costs dominate).  Measured on the
 * 256-instance workload: at 5.5e8 photons 8/16/32/64/128/256 photons per
 * lane give 13.2/16.4/19.7/22.8/25.4/25.8 k photons/ms (256: 11 chunks per
 * wave), at 5.5e7 photons 8/16/32/64 give 13.0/15.9/17.7/18.6 k (64: 4.4
 * chunks per wave).  The workgroup size and the tally cache (no hot-cell
 * box, or the largest that fits the LDS) are then measured on a reduced
 * copy (every instance, photon counts scaled down); the fastest pair wins.
 * Results are cached per batch signature for the process lifetime. */
static const double kChunksPerWave = 4.0;
static std::mutex g_tune_mu;
static std::map<std::pair<int, long long>, BatchShape> g_tune_cache;

static BatchShape autotune(Config** cfgs, int n, int tally, int device, double fraction, hipStream_t s) {
    long long totph = 0;
    for (int i = 0; i < n; i++) totph += (long long)cfgs[i]->nphoton;
    const auto key = std::make_pair(n * 4 + tally, totph);
    {
        std::lock_guard<std::mutex> lk(g_tune_mu);
        auto it = g_tune_cache.find(key);
        if (it != g_tune_cache.end()) return it->second;
    }
    static const int ppl_c[] = {16, 32, 64, 128, 256};
    static const int ppl_small[] = {1, 2, 4, 8, 16};
    static const int blk_c[] = {64, 256};
    int vgprs = 0;
    HIP_CHECK(mcxk_batch_attrs(tally, &vgprs));
    hipDeviceProp_t dp;
    HIP_CHECK(hipGetDeviceProperties(&dp, device));
    const double waves = (double)dp.multiProcessorCount * 4 * waves_per_simd(vgprs);
    BatchShape best;
    best.photons_per_lane = 0;
    best.block = 256;
    for (int ppl : ppl_c) {
        double chunks = 0.0;
        for (int i = 0; i < n; i++) chunks += ceil(ceil((double)cfgs[i]->nphoton / ppl) / 64.0);
        if (chunks >= kChunksPerWave * waves) best.photons_per_lane = ppl;
    }
    double best_rate = -1.0;
    if (best.photons_per_lane == 0) {
        /* a small batch: its launch is bound by the longest chain of
         * histories a lane runs, which fewer photons per lane shorten,
         * against the fixed cost of each chunk, which more of them
         * amortise; such a batch runs in a fraction of a second, so each
         * candidate is timed on the full batch */
        fraction = 1.0;
        for (int ppl : ppl_small) {
            BatchShape t;
            t.photons_per_lane = ppl;
            t.block = 256;
            BatchRun tr(cfgs, n, tally, t, 1.0, device);
            tr.setup();
            float kms = 0.f;
            tr.run(s, false, 1, &kms);
            const double rate = tr.photons / std::max(1e-3, (double)kms);
            if (rate > best_rate) {
                best_rate = rate;
                best.photons_per_lane = ppl;
            }
        }
        best_rate = -1.0;
    }
    /* one reduced copy of the batch (arenas at full size, photon counts
     * scaled), relaunched with each workgroup size and tally cache */
    BatchShape sh = best;
    sh.block = blk_c[0];
    sh.boxcells = 0;
    BatchRun r(cfgs, n, tally, sh, fraction, device);
    r.setup();
    for (int b : blk_c) {
        for (int box : {0, INT_MAX}) {
            r.reshape(b, box);
            float kms = 0.f;
            r.run(s, false, 1, &kms);
            const double rate = r.photons / std::max(1e-3, (double)kms);
            if (rate > best_rate) {
                best_rate = rate;
                best.block = b;
                best.boxcells = box;
            }
        }
    }
    std::lock_guard<std::mutex> lk(g_tune_mu);
    g_tune_cache[key] = best;
    return best;
}

static int run_batch_impl(Config** cfgs, int n, GPUInfo* gpu, McxBatchOptions* opt) {
    if (!cfgs || n <= 0 || !opt) MCX_THROW(-1, "mcx_run_batch: no instances");
    if (n > 65535) MCX_THROW(-1, "mcx_run_batch: at most 65535 instances per batch");
    for (int i = 0; i < n; i++)
        if (!cfgs[i]) MCX_THROW(-1, "mcx_run_batch: null Config pointer");
    const int tally = opt->tally_mode ? opt->tally_mode : MCX_TALLY_F32_HILO;
    if (tally != MCX_TALLY_F32_HILO && tally != MCX_TALLY_F64)
        MCX_THROW(-1, "mcx_run_batch: tally_mode must be 1 (FP32 hi/lo) or 2 (FP64)");
    int device = 0;
    if (cfgs[0]->deviceid[0] > 0) device = cfgs[0]->deviceid[0] - 1;
    (void)gpu;
    const auto t0 = std::chrono::steady_clock::now();
    phase("start", t0);
    HIP_CHECK(hipSetDevice(device));
    Stream st;
    BatchShape shape;
    shape.photons_per_lane = opt->photons_per_lane;
    shape.block = opt->block_size;
    /* tally cache: < 0 none, > 0 at most that many cells per wave, 0 the
     * autotuner's choice (the largest that fits when it does not run) */
    shape.boxcells = opt->tally_cache_cells < 0 ? 0 : (opt->tally_cache_cells > 0 ? opt->tally_cache_cells : INT_MAX);
    bool need_ppl = false;
    for (int i = 0; i < n; i++)
        if (cfgs[i]->autopilot || cfgs[i]->nthread <= 0) need_ppl = true;
    if (shape.photons_per_lane < 0) MCX_THROW(-1, "mcx_run_batch: photons_per_lane must be >= 0");
    if ((need_ppl && shape.photons_per_lane == 0) || shape.block == 0) {
        const double frac = opt->tune_fraction > 0.f ? (double)opt->tune_fraction : 1.0 / 16.0;
        phase("autotune", t0);
        BatchShape t = autotune(cfgs, n, tally, device, frac, st.s);
        phase("autotune done", t0);
        if (shape.photons_per_lane == 0 && need_ppl) shape.photons_per_lane = t.photons_per_lane;
        if (shape.block == 0) shape.block = t.block;
        if (opt->tally_cache_cells == 0) shape.boxcells = t.boxcells;
    }
    if (!need_ppl && opt->photons_per_lane == 0) shape.photons_per_lane = 0;
    BatchRun r(cfgs, n, tally, shape, 1.0, device);
    r.setup();
    phase("setup done", t0);
    float kms = 0.f;
    const float tms = r.run(st.s, opt->use_graph != 0, opt->repeat, &kms);
    phase("run done", t0);
    if (opt->download < 0 || opt->download > 2) MCX_THROW(-1, "mcx_run_batch: download must be 0, 1 or 2");
    const float cw_ms = r.download(st.s, opt->download);
    r.read_counters(st.s, opt->counters, &opt->wave_iterations);
    phase("download done", t0);
    for (int i = 0; i < n; i++) cfgs[i]->runtime = (unsigned int)llround(tms + cw_ms);
    opt->kernel_ms = kms;
    opt->total_ms = tms + cw_ms;
    opt->photons = r.photons;
    opt->chosen_photons_per_lane = shape.photons_per_lane;
    opt->chosen_block_size = shape.block;
    opt->lanes = (double)r.lanes;
    opt->vgprs = r.vgprs;
    opt->chosen_tally_cache_cells = r.box_cells();
    return 0;
}

}  // namespace mcx

using namespace mcx;

extern "C" int mcx_run_batch(Config** cfgs, int ncfg, GPUInfo* gpu, McxBatchOptions* opt) {
    EntryGuard g;
    try {
        return run_batch_impl(cfgs, ncfg, gpu, opt);
    } catch (const Error& e) {
        record_error(e.id, e.msg.c_str());
        return e.id ? e.id : -1;
    }
}
