/*
 * mcx_oracle.cpp -- TEST INFRASTRUCTURE ONLY.
 *
 * CPU restatement of the reference's transport kernel (label and
 * property-word media, non-polarised), one OpenMP iteration per reference
 * virtual thread, each owning its xorshift128+ stream.  Used by tests/ to
 * check the HIP kernel and by bench.py's cpu_baseline leg.  The product
 * never links this file.
 *
 * Followed files (reference, read as text):
 *   src/mcx_core.cu:1767-2528  mcx_main_loop        -> run_thread()
 *   src/mcx_core.cu:1089-1717  launchnewphoton      -> launch()
 *   src/mcx_core.cu:866-956    skipvoid             -> skipvoid()
 *   src/mcx_core.cu:493-507    hitgrid              -> hitgrid()
 *   src/mcx_core.cu:523-537    transmit             -> transmit()
 *   src/mcx_core.cu:552-570    reflectcoeff         -> reflectcoeff()
 *   src/mcx_core.cu:470-478    mcx_nextafterf       -> nextafter_mcx()
 *   src/mcx_core.cu:1011-1031  rotatevector         -> rotatevector()
 *   src/mcx_core.cu:970-999    rotatevector2d       -> rotatevector2d()
 *   src/mcx_core.cu:582-687    updateproperty       -> decode()
 *   src/mcx_core.cu:258-270    finddetector         -> finddetector()
 *   src/mcx_core.cu:365-417    savedetphoton        -> savedet()
 *   src/mcx_core.cu:2191-2292  fluence tally        -> deposit()
 *   src/mcx_rand_xorshift128p.cu:56-124  RNG         -> Rng
 *   src/mcx_core.cu:3247-3251, :3381-3387  host seeding -> mcx_oracle_seedwords()
 *   src/mcx_utils.c:4214-4340  mcx_maskdet          -> mcx_oracle_maskdet()
 *
 * Float arithmetic is written operation by operation (no contraction, see
 * the Makefile) and the elementary functions come from oracle_math.h.
 */
#include "mcx_oracle.h"
#include "oracle_math.h"

#include <algorithm>
#include <atomic>
#include <climits>
#include <utility>
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

const float EPSF = 1.1920928955078125e-07f; /* FLT_EPSILON */
const unsigned int DET_MASK = 0x80000000u;
const unsigned int MED_MASK = 0x7FFFFFFFu;
const unsigned int OUT_MIN = 0xFFFFFFFFu;
const unsigned int OUT_MAX = 0x7FFFFFFFu;
const float TWO_PI = 6.28318530717959f;
const float MAX_ACCUM = 1000.f;
const float ROULETTE = 10.f;
constexpr float JUST_BELOW_ONE = 0.9998f;
/* the engine's per-photon step limit (an internal-error exit, never taken
 * by a physical history with nanosecond time windows) */
constexpr unsigned long long MAX_PHOTON_STEPS = 1ull << 26;
enum { bcUnknown, bcReflect, bcAbsorb, bcMirror, bcCyclic };
enum { otFlux, otFluence, otEnergy, otJacobian, otWP, otDCS, otRF, otL, otRFmus, otWLTOF, otWPTOF };

/* Planted mistakes for the tests that prove the pins catch them (0 = the
 * faithful restatement; the ids are listed in mcx_oracle.h).  Set only
 * between runs, never during one. */
int g_mut = 0;

/* The RF replay's modulation phasor sin, cos of x = omega * tof: the loop
 * and mcx_oracle_sincosf both call this.  Mutation 21 drops the last
 * part of the pi/2 reduction. */
inline void rf_sincos(float x, float* s, float* c) {
    if (g_mut != 21) {
        om_sincosf(x, s, c);
        return;
    }
    const float q = rintf(x * 0.636619772367581343f);
    float r = fmaf(-q, 1.5703125f, x);
    r = fmaf(-q, 4.837512969970703125e-04f, r);
    om_sincos_series(r, ((int)q) & 3, s, c);
}

/* ---- glibc rand(): TYPE_3 additive feedback, r[i] = r[i-3] + r[i-31] --- */
struct GlibcRand {
    int32_t r[34];
    uint32_t ring[31];
    int pos;
    explicit GlibcRand(uint32_t seed) {
        if (seed == 0) seed = 1;
        std::vector<uint32_t> st(400);
        int32_t w = (int32_t)seed;
        st[0] = (uint32_t)w;
        for (int i = 1; i < 31; i++) {
            /* 16807 * w mod (2^31-1) by Schrage's decomposition */
            int32_t hi = w / 127773, lo = w % 127773;
            w = 16807 * lo - 2836 * hi;
            if (w < 0) w += 2147483647;
            st[i] = (uint32_t)w;
        }
        for (int i = 31; i < 34; i++) st[i] = st[i - 31];
        for (int i = 34; i < 344; i++) st[i] = st[i - 31] + st[i - 3];
        for (int i = 0; i < 31; i++) ring[i] = st[313 + i]; /* r[313..343] */
        pos = 0;
    }
    /* next output: r[i] = r[i-31] + r[i-3]; ring holds r[i-31..i-1] */
    int next() {
        uint32_t v = ring[pos] + ring[(pos + 28) % 31];
        ring[pos] = v;
        pos = (pos + 1) % 31;
        return (int)(v >> 1);
    }
};

/* ---- xorshift128+ (src/mcx_rand_xorshift128p.cu:56-72) ----------------- */
struct Rng {
    uint64_t t0, t1;
    void seed(const unsigned int* s) {
        t0 = ((uint64_t)s[0] << 32) | s[1];
        t1 = ((uint64_t)s[2] << 32) | s[3];
    }
    /* the state as the 4 words of u64 {s0, s1} in little-endian memory order */
    void setstate(const unsigned int* w) {
        t0 = ((uint64_t)w[1] << 32) | w[0];
        t1 = ((uint64_t)w[3] << 32) | w[2];
    }
    void getstate(unsigned int* w) const {
        w[0] = (unsigned int)t0; w[1] = (unsigned int)(t0 >> 32);
        w[2] = (unsigned int)t1; w[3] = (unsigned int)(t1 >> 32);
    }
    float next() {
        uint64_t s1 = t0;
        const uint64_t s0 = t1;
        t0 = s0;
        s1 ^= s1 << 23;
        t1 = s1 ^ s0 ^ (s1 >> 18) ^ (s0 >> 5);
        uint32_t lo = (uint32_t)(t1 + s0);
        return om_bits2f(0x3F800000u | (lo >> 9)) - 1.0f;
    }
    float scatlen() { return -om_logf(next() + EPSF); }
};

struct Vec4 { float x, y, z, w; };

struct Ctx {
    const OracleParams* P;
    OracleOut* O;
    unsigned int dimlen_x, dimlen_y, dimlen_z;
    float maxidx_x, maxidx_y, maxidx_z;
    unsigned int maxmedia, partialdata, w0offset, reclen;
    size_t shadowoff;     /* FP32 tally: offset of the MAX_ACCUM shadow half */
    int rfshift;          /* RF replay: gates from the real plane to the imaginary one; 0 otherwise */
    std::atomic<unsigned int>* detected;
    const Vec4* gprop;   /* media then detectors */
};

inline float hitgrid(const Vec4& p, const Vec4& v, const float* rv, int* id) {
    const int m = (g_mut == 4);
    float h0 = fabsf(((float)(id[0] + (m || v.x > 0.f)) - p.x) * rv[0]);
    float h1 = fabsf(((float)(id[1] + (m || v.y > 0.f)) - p.y) * rv[1]);
    float h2 = fabsf(((float)(id[2] + (m || v.z > 0.f)) - p.z) * rv[2]);
    float dist = fminf(fminf(h0, h1), h2);
    id[3] = (dist == h0 ? 0 : (dist == h1 ? 1 : 2));
    return dist;
}

inline void rotatevector(Vec4& v, float stheta, float ctheta, float sphi, float cphi) {
    if (v.z > -1.f + EPSF && v.z < 1.f - EPSF) {
        float tmp0 = 1.f - v.z * v.z;
        float tmp1 = stheta / sqrtf(tmp0);
        float nx = tmp1 * (v.x * v.z * cphi - v.y * sphi) + v.x * ctheta;
        float ny = tmp1 * (v.y * v.z * cphi + v.x * sphi) + v.y * ctheta;
        float nz = -tmp1 * tmp0 * cphi + v.z * ctheta;
        v.x = nx; v.y = ny; v.z = nz;
    } else {
        float nz = (v.z > 0.f) ? ctheta : -ctheta;
        v.x = stheta * cphi;
        v.y = stheta * sphi;
        v.z = nz;
    }
    float tmp0 = 1.f / sqrtf(v.x * v.x + v.y * v.y + v.z * v.z);
    v.x *= tmp0; v.y *= tmp0; v.z *= tmp0;
}

inline void rotatevector2d(Vec4& v, int is2d, float stheta, float ctheta) {
    if (is2d == 1) {
        float ny = v.y * ctheta - v.z * stheta, nz = v.y * stheta + v.z * ctheta;
        v.x = 0.f; v.y = ny; v.z = nz;
    } else if (is2d == 2) {
        float nx = v.x * ctheta - v.z * stheta, nz = v.x * stheta + v.z * ctheta;
        v.x = nx; v.y = 0.f; v.z = nz;
    } else if (is2d == 3) {
        float nx = v.x * ctheta - v.y * stheta, ny = v.x * stheta + v.y * ctheta;
        v.x = nx; v.y = ny; v.z = 0.f;
    }
    float tmp0 = 1.f / sqrtf(v.x * v.x + v.y * v.y + v.z * v.z);
    v.x *= tmp0; v.y *= tmp0; v.z *= tmp0;
}

inline float comp(const Vec4& v, int i) { return i == 0 ? v.x : (i == 1 ? v.y : v.z); }
inline float sgnf(float x) { return (float)((x > 0.f) - (x < 0.f)); }

inline void transmit(Vec4& v, float n1, float n2, int flipdir) {
    float tmp0 = (g_mut == 5) ? n2 / n1 : n1 / n2;
    v.x *= tmp0; v.y *= tmp0; v.z *= tmp0;
    if (flipdir == 0) {
        tmp0 = v.y * v.y + v.z * v.z;
        v.x = (tmp0 < 1.f) ? sqrtf(1.f - tmp0) * sgnf(v.x) : 0.f;
    } else if (flipdir == 1) {
        tmp0 = v.x * v.x + v.z * v.z;
        v.y = (tmp0 < 1.f) ? sqrtf(1.f - tmp0) * sgnf(v.y) : 0.f;
    } else {
        tmp0 = v.x * v.x + v.y * v.y;
        v.z = (tmp0 < 1.f) ? sqrtf(1.f - tmp0) * sgnf(v.z) : 0.f;
    }
    tmp0 = 1.f / sqrtf(v.x * v.x + v.y * v.y + v.z * v.z);
    v.x *= tmp0; v.y *= tmp0; v.z *= tmp0;
}

inline float reflectcoeff(const Vec4& v, float n1, float n2, int flipdir) {
    if (g_mut == 6) std::swap(n1, n2);
    float Icos = fabsf(comp(v, flipdir));
    float tmp0 = n1 * n1;
    float tmp1 = n2 * n2;
    float tmp2 = 1.f - tmp0 / tmp1 * (1.f - Icos * Icos);
    if (tmp2 > 0.f) {
        float Re = tmp0 * Icos * Icos + tmp1 * tmp2;
        tmp2 = sqrtf(tmp2);
        float Im = 2.f * n1 * n2 * Icos * tmp2;
        float Rtotal = (Re - Im) / (Re + Im);
        Re = tmp1 * Icos * Icos + tmp0 * tmp2 * tmp2;
        Rtotal = (Rtotal + (Re - Im) / (Re + Im)) * 0.5f;
        return Rtotal;
    }
    return 1.f;
}

/* src/mcx_core.cu:1949-1957: Henyey-Greenstein inversion of u, clamped */
inline float hg_costheta(float g, float u) {
    float tmp0 = (1.f - g * g) / (1.f - g + 2.f * g * u);
    if (g_mut != 2) tmp0 *= tmp0;
    tmp0 = (1.f + g * g - tmp0) / (2.f * g);
    if (g_mut == 1) tmp0 = -tmp0;
    return fmaxf(-1.f, fminf(1.f, tmp0));
}

/* src/mcx_core.cu:1946-1964: scattering angle cosine from one uniform u:
 * HG with g when the selector gsel (g, or 0 after gscatter events) is
 * beyond EPS, isotropic otherwise */
inline float scatter_costheta(float gsel, float g, float u) {
    return (fabsf(gsel) > EPSF) ? hg_costheta(g, u) : 2.f * u - 1.f;
}

/* src/mcx_core.cu:2418-2432: unpolarised Fresnel reflectance at a voxel
 * face for |cos| of the incidence angle cphi; 1 beyond the critical angle.
 * len = cos^2 of the refraction angle, (ctheta -/+ stheta) the s and then
 * p amplitudes' numerator/denominator. */
inline float fresnel_face(float n1, float n2, float cphi) {
    const float tmp0 = n1 * n1;
    const float tmp1 = n2 * n2;
    const float sphi = 1.f - cphi * cphi;
    const float len = 1.f - tmp0 / tmp1 * sphi;
    if (!(len > 0.f)) return 1.f;
    float ctheta = tmp0 * cphi * cphi + tmp1 * len;
    const float stheta = 2.f * n1 * n2 * cphi * sqrtf(len);
    float Rtotal = (ctheta - stheta) / (ctheta + stheta);
    if (g_mut != 3) ctheta = tmp1 * cphi * cphi + tmp0 * len;
    Rtotal = (Rtotal + (ctheta - stheta) / (ctheta + stheta)) * 0.5f;
    return Rtotal;
}

/* src/mcx_core.cu:2127-2132: move by len and, unless the step ends inside
 * the voxel (scattering site or interface plane), enter the neighbour
 * across face fd[3] */
inline void advance(Vec4& p, const Vec4& v, float len, bool stay, int* fd) {
    p.x = p.x + len * v.x; p.y = p.y + len * v.y; p.z = p.z + len * v.z;
    if (fd[3] == 0) fd[0] += stay ? 0 : (v.x > 0.f ? 1 : -1);
    if (fd[3] == 1) fd[1] += stay ? 0 : (v.y > 0.f ? 1 : -1);
    if (fd[3] == 2) fd[2] += stay ? 0 : (v.z > 0.f ? 1 : -1);
}

inline float nextafter_mcx(float a, int dir, int maxvoidstep) {
    float f = a + (float)maxvoidstep;
    uint32_t i = om_f2bits(f);
    i += (uint32_t)dir ^ (i & 0x80000000u);
    return om_bits2f(i) - (float)maxvoidstep;
}

inline void atomic_add_double(double* a, double v) {
#pragma omp atomic
    *a += v;
}

inline float atomic_add_float_ret(float* a, float v) {
    std::atomic_ref<float> r(*a);
    return r.fetch_add(v);
}

/* One virtual thread's state (the reference keeps these in registers). */
struct Thread {
    Ctx* C;
    unsigned int idx;
    int nphoton;          /* threadphoton + (idx < oddphotons) */
    Vec4 p, v, f;         /* f = {pscat, t, pathlen, ndone} */
    float rv[3];
    Vec4 prop;
    Rng t;
    int flipdir[4];
    unsigned int idx1d, mediaid;
    float w0;
    float* ppath;
    unsigned int gid = 0;        /* global index of the current photon */
    unsigned int seedsave[4];    /* launch RNG state of the current photon */
    unsigned long long nstep = 0, ndep = 0, nscat = 0, nlaunch = 0, pstep = 0;
    int err = 0;
    /* split-voxel state (src/mcx_core.h:65-83) */
    float rp[3] = {0.f, 0.f, 0.f}, nv[3] = {0.f, 0.f, 0.f};
    unsigned int svlo = 0, svup = 0, svupper = 0, svsplit = 0;
    bool testint = false, hitintf = false;

    const OracleParams& P() const { return *C->P; }

    unsigned int finddetector() const {
        const Ctx& c = *C;
        for (unsigned int i = c.maxmedia + 1; i < c.maxmedia + P().detnum + 1; i++) {
            const Vec4& d = c.gprop[i];
            if ((d.x - p.x) * (d.x - p.x) + (d.y - p.y) * (d.y - p.y) + (d.z - p.z) * (d.z - p.z) < d.w * d.w)
                return (d.w < 0.f) ? 0 : i - c.maxmedia;
        }
        return 0;
    }

    void savedet(unsigned int isdet) {
        int detid = (isdet == OUT_MIN) ? -1 : (int)finddetector();
        if (!detid) return;
        unsigned int base = C->detected->fetch_add(1);
        if (base < P().maxdetphoton) {
            if (!C->O->detdata) return;
            float* rec = C->O->detdata + (size_t)base * C->reclen;
            unsigned int k = 0, flag = P().savedetflag;
            if (flag & 1) rec[k++] = (float)detid;
            for (unsigned int i = 0; i < C->partialdata; i++) rec[k++] = ppath[i];
            if ((flag >> 4) & 1) { rec[k++] = p.x; rec[k++] = p.y; rec[k++] = p.z; }
            if ((flag >> 5) & 1) { rec[k++] = v.x; rec[k++] = v.y; rec[k++] = v.z; }
            if ((flag >> 6) & 1) rec[k++] = ppath[C->w0offset - 2];
            if (P().saveseed && C->O->seeddata)
                for (int i = 0; i < 4; i++) C->O->seeddata[(size_t)base * 4 + i] = seedsave[i];
        } else if (P().savedet == 3) {
            C->detected->fetch_sub(1);
        }
    }

    /* IEEE binary16 bits -> float, exact */
    static float half_to_float(unsigned int h) {
        const unsigned int e = (h >> 10) & 0x1Fu, m = h & 0x3FFu;
        float v;
        if (e == 0)
            v = ldexpf((float)m, -24);
        else if (e == 31)
            v = m ? NAN : INFINITY;
        else
            v = ldexpf((float)(m | 0x400u), (int)e - 25);
        return ((h >> 15) & 1u) ? -v : v;
    }

    bool words() const { return P().mediaformat > 4 && P().mediaformat != 97; }
    bool split() const { return P().mediaformat == 97; }
    unsigned int sv_label() const { return svupper ? svup : svlo; }

    /* src/mcx_core.cu:688-729 (updateproperty, split voxels) */
    void svmc_update(Vec4& out, unsigned int w, unsigned int vox, const Vec4& q, int i0, int i1, int i2) {
        const Ctx& c = *C;
        if (vox == OUT_MIN || vox == OUT_MAX) {
            out = c.gprop[0];
            return;
        }
        const unsigned int lo = P().media[(size_t)vox + (size_t)P().nx * P().ny * P().nz];
        const unsigned int hi = w & MED_MASK;
        svlo = hi >> 24;
        svup = (hi >> 16) & 0xFFu;
        if (svup) {
            rp[0] = (float)((hi >> 8) & 0xFFu) * (1.f / 255.f);
            rp[1] = (float)(hi & 0xFFu) * (1.f / 255.f);
            rp[2] = (float)(lo >> 24) * (1.f / 255.f);
            rp[0] += (float)i0;
            rp[1] += (float)i1;
            rp[2] += (float)i2;
            nv[0] = (float)((lo >> 16) & 0xFFu) * (2.f / 255.f) - 1.f;
            nv[1] = (float)((lo >> 8) & 0xFFu) * (2.f / 255.f) - 1.f;
            nv[2] = (float)(lo & 0xFFu) * (2.f / 255.f) - 1.f;
            const float r = 1.f / sqrtf(nv[0] * nv[0] + nv[1] * nv[1] + nv[2] * nv[2]);
            for (int k = 0; k < 3; k++) nv[k] = nv[k] * r;
            if ((rp[0] - q.x) * nv[0] + (rp[1] - q.y) * nv[1] + (rp[2] - q.z) * nv[2] < 0.f) {
                out = c.gprop[svup];
                svupper = 1;
                for (int k = 0; k < 3; k++) nv[k] = -nv[k];
            } else {
                out = c.gprop[svlo];
                svupper = 0;
            }
            svsplit = 1;
        } else {
            out = c.gprop[svlo];
            svsplit = 0;
            svupper = 0;
        }
    }

    /* src/mcx_core.cu:747-766 (ray_plane_intersect) */
    bool ray_plane(float& len, float& slen) const {
        if (v.x * nv[0] + v.y * nv[1] + v.z * nv[2] <= 0.f) return false;
        const float p1x = p.x + len * v.x, p1y = p.y + len * v.y, p1z = p.z + len * v.z;
        const float d0 = (p.x - rp[0]) * nv[0] + (p.y - rp[1]) * nv[1] + (p.z - rp[2]) * nv[2];
        const float d1 = (p1x - rp[0]) * nv[0] + (p1y - rp[1]) * nv[1] + (p1z - rp[2]) * nv[2];
        if (d0 * d1 > 0.f) return false;
        const float len0 = len * d0 / (d0 - d1);
        len = (len0 > 0.f) ? len0 : len;
        slen = len * prop.y * ((v.w + 1.f > (float)P().gscatter) ? (1.f - prop.z) : 1.f);
        return true;
    }

    /* src/mcx_core.cu:782-828 (reflectray); true: transmitted into the background */
    bool reflectray(float n1) {
        const Ctx& c = *C;
        const float Icos = fabsf(v.x * nv[0] + v.y * nv[1] + v.z * nv[2]);
        const float n2 = c.gprop[svupper ? svup : svlo].w;
        const float tmp0 = n1 * n1, tmp1 = n2 * n2;
        float tmp2 = 1.f - tmp0 / tmp1 * (1.f - Icos * Icos);
        float* vv[3] = {&v.x, &v.y, &v.z};
        if (tmp2 > 0.f) {
            float Re = tmp0 * Icos * Icos + tmp1 * tmp2;
            tmp2 = sqrtf(tmp2);
            const float Im = 2.f * n1 * n2 * Icos * tmp2;
            float Rtotal = (Re - Im) / (Re + Im);
            Re = tmp1 * Icos * Icos + tmp0 * tmp2 * tmp2;
            Rtotal = (Rtotal + (Re - Im) / (Re + Im)) * 0.5f;
            if (t.next() <= Rtotal) {
                for (int k = 0; k < 3; k++) *vv[k] = *vv[k] + (-2.f * Icos) * nv[k];
                svupper = !svupper;
            } else {
                for (int k = 0; k < 3; k++) *vv[k] = *vv[k] + (-Icos) * nv[k];
                const float r = n1 / n2;
                for (int k = 0; k < 3; k++) *vv[k] = tmp2 * nv[k] + r * *vv[k];
                for (int k = 0; k < 3; k++) nv[k] = -nv[k];
                if ((svupper ? svupper : svlo) == 0) return true;
                prop = c.gprop[svupper ? svup : svlo];
            }
        } else {
            for (int k = 0; k < 3; k++) *vv[k] = *vv[k] + (-2.f * Icos) * nv[k];
            svupper = !svupper;
        }
        const float r = 1.f / sqrtf(v.x * v.x + v.y * v.y + v.z * v.z);
        v.x = v.x * r;
        v.y = v.y * r;
        v.z = v.z * r;
        rv[0] = 1.f / v.x; rv[1] = 1.f / v.y; rv[2] = 1.f / v.z;
        return false;
    }

    /* src/mcx_core.cu:582-687 (updateproperty, non-label formats): fields a
     * format does not encode keep their value in q; the mixed-label format
     * draws one number and narrows m to the label drawn */
    void decode(Vec4& q, unsigned int& m, unsigned int vox) {
        const Ctx& c = *C;
        const unsigned int w = m & MED_MASK;
        const int fmt = P().mediaformat;
        if (fmt == 99) {                       /* label + one half-precision property */
            q = c.gprop[w & 0x3FFFu];
            const float h = fabsf(half_to_float(w >> 16));
            switch ((w >> 14) & 3u) {
                case 0: q.x = h; break;
                case 1: q.y = h; break;
                case 2: q.z = h; break;
                default: q.w = h;
            }
        } else if (fmt == 101) {               /* float mua */
            float f;
            memcpy(&f, &m, 4);
            q.x = fabsf(f);
            q.w = c.gprop[w != 0u ? 1 : 0].w;
        } else if (fmt == 100 || fmt == 102) { /* half mua, mus */
            q.x = fabsf(half_to_float(w & 0xFFFFu));
            q.y = fabsf(half_to_float(w >> 16));
            q.w = c.gprop[w != 0u ? 1 : 0].w;
        } else if (fmt == 96) {                /* half mua, mus, g, n in two words */
            if (vox == OUT_MIN || vox == OUT_MAX) {
                q = c.gprop[0];
                return;
            }
            const unsigned int w2 = P().media[(size_t)vox + (size_t)P().nx * P().ny * P().nz];
            q.x = fabsf(half_to_float(w & 0xFFFFu));
            q.y = fabsf(half_to_float(w >> 16));
            q.z = fabsf(half_to_float(w2 & 0xFFFFu));
            q.w = fabsf(half_to_float(w2 >> 16));
        } else if (fmt == 98) {                /* two labels mixed by a fraction */
            const unsigned int frac = w >> 16;
            if (frac > 0u) {
                if (t.next() * 32767.f < (float)frac) {
                    q = c.gprop[(w >> 8) & 0xFFu];
                    m >>= 8;
                } else {
                    q = c.gprop[w & 0xFFu];
                }
                m &= 0xFFFFu;
            } else {
                q = c.gprop[w & 0xFFu];
            }
        } else if (fmt == 103) {               /* byte fractions between rows 1 and 2 */
            const Vec4 a = c.gprop[1], b = c.gprop[2];
            q.x = (float)(w & 0xFFu) * (1.f / 255.f) * (b.x - a.x) + a.x;
            q.y = (float)((w >> 8) & 0xFFu) * (1.f / 255.f) * (b.y - a.y) + a.y;
            q.z = (float)((w >> 16) & 0xFFu) * (1.f / 255.f) * (b.z - a.z) + a.z;
            q.w = (float)(w >> 24) * (1.f / 127.f) * (b.w - a.w) + a.w;
        } else if (fmt == 104) {               /* short fractions between rows 1 and 2 */
            const Vec4 a = c.gprop[1], b = c.gprop[2];
            q.x = (float)(w & 0xFFFFu) * (1.f / 65535.f) * (b.x - a.x) + a.x;
            q.y = (float)(w >> 16) * (1.f / 65535.f) * (b.y - a.y) + a.y;
            q.w = c.gprop[w != 0u ? 1 : 0].w;
        }
    }

    void setprop(unsigned int& m, unsigned int vox) {
        if (split())
            svmc_update(prop, m, vox, p, flipdir[0], flipdir[1], flipdir[2]);
        else if (words())
            decode(prop, m, vox);
        else
            prop = C->gprop[m & MED_MASK];
    }

    /* src/mcx_core.cu:866-956 */
    int skipvoid() {
        const OracleParams& Q = P();
        int count = 1, id[4] = {0, 0, 0, -1};
        id[0] = (int)floorf(p.x); id[1] = (int)floorf(p.y); id[2] = (int)floorf(p.z);
        while (1) {
            if ((unsigned)id[0] < Q.nx && (unsigned)id[1] < Q.ny && (unsigned)id[2] < Q.nz) {
                unsigned int i1 = id[2] * C->dimlen_y + id[1] * C->dimlen_x + id[0];
                if (Q.media[i1] & MED_MASK) {
                    p.x -= v.x; p.y -= v.y; p.z -= v.z;
                    id[0] = (int)floorf(p.x); id[1] = (int)floorf(p.y); id[2] = (int)floorf(p.z);
                    f.y -= Q.minaccumtime;
                    i1 = id[2] * C->dimlen_y + id[1] * C->dimlen_x + id[0];
                    count = 0;
                    while (!((unsigned)id[0] < Q.nx && (unsigned)id[1] < Q.ny && (unsigned)id[2] < Q.nz) ||
                           !(Q.media[i1] & MED_MASK)) {
                        float dist = hitgrid(p, v, rv, id);
                        f.y += Q.minaccumtime * dist;
                        p.x = p.x + dist * v.x; p.y = p.y + dist * v.y; p.z = p.z + dist * v.z;
                        if (id[3] == 0) id[0] += (v.x > 0.f ? 1 : -1);
                        if (id[3] == 1) id[1] += (v.y > 0.f ? 1 : -1);
                        if (id[3] == 2) id[2] += (v.z > 0.f ? 1 : -1);
                        i1 = id[2] * C->dimlen_y + id[1] * C->dimlen_x + id[0];
                        if (count++ > 3) break;
                    }
                    /* entry search gave up outside the grid: no tissue entry */
                    if (!((unsigned)id[0] < Q.nx && (unsigned)id[1] < Q.ny && (unsigned)id[2] < Q.nz)) return -1;
                    f.y = Q.voidtime ? f.y : 0.f;
                    Vec4 h;
                    if (split()) {
                        svmc_update(h, Q.media[i1], i1, p, id[0], id[1], id[2]);
                    } else if (words()) {
                        /* src/mcx_core.cu:921-922, decoded into a local */
                        unsigned int w = Q.media[i1];
                        h = C->gprop[1];
                        decode(h, w, i1);
                    } else {
                        h = C->gprop[Q.media[i1] & MED_MASK];
                    }
                    if (Q.isspecular && h.w != C->gprop[0].w) {
                        p.w *= 1.f - reflectcoeff(v, C->gprop[0].w, h.w, id[3]);
                        if (p.w > EPSF) transmit(v, C->gprop[0].w, h.w, id[3]);
                    }
                    return (int)i1;
                }
            }
            if ((p.x < 0.f && v.x <= 0.f) || (p.x >= C->maxidx_x && v.x >= 0.f) ||
                (p.y < 0.f && v.y <= 0.f) || (p.y >= C->maxidx_y && v.y >= 0.f) ||
                (p.z < 0.f && v.z <= 0.f) || (p.z >= C->maxidx_z && v.z >= 0.f))
                return -1;
            p.x = p.x + v.x; p.y = p.y + v.y; p.z = p.z + v.z;
            id[0] = (int)floorf(p.x); id[1] = (int)floorf(p.y); id[2] = (int)floorf(p.z);
            f.y += Q.minaccumtime;
            if (count++ > Q.maxvoidstep) return -1;
        }
    }

    void locate_launch() {
        const OracleParams& Q = P();
        idx1d = (unsigned)((int)floorf(p.z) * (int)C->dimlen_y + (int)floorf(p.y) * (int)C->dimlen_x + (int)floorf(p.x));
        if (p.x < 0.f || p.y < 0.f || p.z < 0.f || p.x >= C->maxidx_x || p.y >= C->maxidx_y || p.z >= C->maxidx_z)
            mediaid = 0;
        else
            mediaid = Q.media[idx1d];
    }

    /* src/mcx_core.cu:1089-1717; returns 0 launched, 1 done, -1 failed */
    int launch(unsigned int isdet) {
        const OracleParams& Q = P();
        w0 = 1.f;
        int canfocus = 1;
        if (Q.savedet == 3 && C->detected->load() >= Q.maxdetphoton) return 1;
        if (fabsf(p.w) >= 0.f) {
            ppath[C->partialdata] += p.w;
            if (Q.issaveref && mediaid == 0 && idx1d != OUT_MIN && idx1d != OUT_MAX && p.w > 0.f) {
                int ts = std::max(0, std::min((int)Q.maxgate - 1, (int)floorf((f.y - Q.twin0) * Q.Rtstep)));
                deposit(idx1d, ts, -p.w);
            }
            if (Q.savedet && (isdet & DET_MASK) == DET_MASK && (mediaid == 0 || (split() && sv_label() == 0))) savedet(isdet);
        }
        if (Q.savedet)
            for (unsigned int i = 0; i < C->partialdata; i++) ppath[i] = 0.f;
        if ((int)f.w >= nphoton - 1) return 1;
        /* src/mcx_core.cu:1205-1215, with the photon index counted from the
         * thread's first photon idx*threadphoton + min(idx, oddphotons), or
         * (replayindex 1) from idx*threadphoton + min(idx, oddphotons - 1) */
        gid = idx * (unsigned)Q.threadphoton + (Q.replayindex ? (unsigned)std::min((int)idx, Q.oddphotons - 1)
                                                              : std::min(idx, (unsigned)Q.oddphotons)) +
              (unsigned)((int)f.w + 1);
        if (Q.replayseed) t.setstate(Q.replayseed + (size_t)gid * 4);
        if (Q.saveseed) t.getstate(seedsave);
        float* pp = ppath + C->partialdata;
        const bool ispencil = (Q.srctype == 0);
        do {
            p = {Q.srcpos[0], Q.srcpos[1], Q.srcpos[2], Q.srcpos[3]};
            v = {Q.srcdir[0], Q.srcdir[1], Q.srcdir[2], Q.srcdir[3]};
            f = {0.f, 0.f, Q.minaccumtime, f.w};
            memcpy(&idx1d, &Q.srcparam2[2], 4);
            memcpy(&mediaid, &Q.srcparam2[3], 4);
            if (split()) svsplit = svlo = svup = svupper = 0;
            float org[3] = {Q.srcpos[0], Q.srcpos[1], Q.srcpos[2]};
            if (!ispencil) {
                const float* s1 = Q.srcparam1;
                const float* s2 = Q.srcparam2;
                switch (Q.srctype) {
                    case 4: /* planar */
                    case 5: /* pattern */
                    case 6: /* fourier */
                    case 14: /* pencilarray */
                    case 15: { /* pattern3d */
                        float rx = t.next();
                        float ry = t.next();
                        if (Q.srctype == 15) {
                            float rz = t.next();
                            p.x = p.x + rx * s1[0];
                            p.y = p.y + ry * s1[1];
                            p.z = p.z + rz * s1[2];
                            const float ix = (g_mut == 17) ? ry : rx, iy = (g_mut == 17) ? rx : ry;
                            p.w = Q.srcpos[3] * Q.srcpattern[(int)(rz * JUST_BELOW_ONE * s1[2]) * (int)s1[1] * (int)s1[0] +
                                                             (int)(iy * JUST_BELOW_ONE * s1[1]) * (int)s1[0] +
                                                             (int)(ix * JUST_BELOW_ONE * s1[0])];
                        } else {
                            p.x = p.x + rx * s1[0] + ry * s2[0];
                            p.y = p.y + rx * s1[1] + ry * s2[1];
                            p.z = p.z + rx * s1[2] + ry * s2[2];
                        }
                        if (Q.srctype == 5) {
                            p.w = Q.srcpos[3] * Q.srcpattern[(int)(ry * JUST_BELOW_ONE * s2[3]) * (int)s1[3] +
                                                             (int)(rx * JUST_BELOW_ONE * s1[3])];
                        } else if (Q.srctype == 6) {
                            float s, c;
                            float ang = (floorf(s1[3]) * rx + floorf(s2[3]) * ry + s1[3] - floorf(s1[3]));
                            om_sincos2pi(ang, &s, &c);
                            p.w = Q.srcpos[3] * (c * (1.f - s2[3] + floorf(s2[3])) + 1.f) * 0.5f;
                        } else if (Q.srctype == 14) {
                            p.x = Q.srcpos[0] + floorf(rx * s1[3]) * s1[0] / (s1[3] - 1.f) + floorf(ry * s2[3]) * s2[0] / (s2[3] - 1.f);
                            p.y = Q.srcpos[1] + floorf(rx * s1[3]) * s1[1] / (s1[3] - 1.f) + floorf(ry * s2[3]) * s2[1] / (s2[3] - 1.f);
                            p.z = Q.srcpos[2] + floorf(rx * s1[3]) * s1[2] / (s1[3] - 1.f) + floorf(ry * s2[3]) * s2[2] / (s2[3] - 1.f);
                        }
                        locate_launch();
                        org[0] = org[0] + (s1[0] + s2[0]) * 0.5f;
                        org[1] = org[1] + (s1[1] + s2[1]) * 0.5f;
                        org[2] = org[2] + (s1[2] + s2[2]) * 0.5f;
                        break;
                    }
                    case 9:    /* fourierx */
                    case 10: { /* fourierx2d */
                        float rx = t.next();
                        float ry = t.next();
                        float v2w = s1[3] * (1.f / sqrtf(s1[0] * s1[0] + s1[1] * s1[1] + s1[2] * s1[2]));
                        float v2x = v2w * (Q.srcdir[1] * s1[2] - Q.srcdir[2] * s1[1]);
                        float v2y = v2w * (Q.srcdir[2] * s1[0] - Q.srcdir[0] * s1[2]);
                        float v2z = v2w * (Q.srcdir[0] * s1[1] - Q.srcdir[1] * s1[0]);
                        if (g_mut == 11) { v2x = -v2x; v2y = -v2y; v2z = -v2z; }
                        p.x = p.x + rx * s1[0] + ry * v2x;
                        p.y = p.y + rx * s1[1] + ry * v2y;
                        p.z = p.z + rx * s1[2] + ry * v2z;
                        if (Q.srctype == 10) {
                            float sa, ca, sb, cb;
                            const float phx = (g_mut == 12) ? s2[3] : s2[2], phy = (g_mut == 12) ? s2[2] : s2[3];
                            om_sincos2pi(s2[0] * rx + phx, &sa, &ca);
                            om_sincos2pi(s2[1] * ry + phy, &sb, &cb);
                            p.w = Q.srcpos[3] * (sa * sb + 1.f) * 0.5f;
                        } else {
                            float sa, ca;
                            om_sincos2pi(s2[0] * rx + s2[1] * ry + s2[2], &sa, &ca);
                            p.w = Q.srcpos[3] * (ca * (1.f - s2[3]) + 1.f) * 0.5f;
                        }
                        locate_launch();
                        org[0] = org[0] + (s1[0] + v2x) * 0.5f;
                        org[1] = org[1] + (s1[1] + v2y) * 0.5f;
                        org[2] = org[2] + (s1[2] + v2z) * 0.5f;
                        break;
                    }
                    case 11: { /* zgaussian */
                        float sphi, cphi, stheta, ctheta;
                        om_sincos2pi(t.next(), &sphi, &cphi);
                        float ang = sqrtf((g_mut == 13 ? 1.f : 2.f) * t.scatlen()) * (1.f - 2.f * t.next()) * s1[0];
                        om_sincos2pi(ang * 0.159154943091895336f, &stheta, &ctheta);
                        rotatevector(v, stheta, ctheta, sphi, cphi);
                        canfocus = 0;
                        break;
                    }
                    case 16: { /* hyperboloid gaussian */
                        float sphi, cphi;
                        om_sincos2pi(t.next(), &sphi, &cphi);
                        float r = sqrtf(0.5f * t.scatlen()) * s1[0];
                        float ax = (g_mut == 16) ? s1[1] / s1[2] : -s1[1] / s1[2];
                        float ay = 1.f / sqrtf(r * r + s1[2] * s1[2]);
                        float px = r * (cphi - ax * sphi), py = r * (sphi + ax * cphi), pz = 0.f;
                        float dx = -r * sphi * ay, dy = r * cphi * ay, dz = s1[2] * ay;
                        if (v.z > -1.f + EPSF && v.z < 1.f - EPSF) {
                            r = 1.f - v.z * v.z;
                            float st = sqrtf(r);
                            r = 1.f / sqrtf(r);
                            float cp = v.x * r, sp = v.y * r;
                            float qx = px * cp * v.z - py * sp, qy = px * sp * v.z + py * cp, qz = -px * st;
                            px = qx; py = qy; pz = qz;
                            float nx = dx * cp * v.z - dy * sp + dz * cp * st;
                            float ny = dx * sp * v.z + dy * cp + dz * sp * st;
                            float nz = -dx * st + dz * v.z;
                            v.x = nx; v.y = ny; v.z = nz;
                        } else {
                            v.z = (v.z > 0.f) ? dz : -dz;
                            v.x = dx; v.y = dy;
                        }
                        p.x = px + Q.srcpos[0];
                        p.y = py + Q.srcpos[1];
                        p.z = pz + Q.srcpos[2];
                        locate_launch();
                        canfocus = 0;
                        break;
                    }
                    case 12:   /* line */
                    case 13: { /* slit */
                        float r = t.next();
                        p.x = p.x + r * s1[0];
                        p.y = p.y + r * s1[1];
                        p.z = p.z + r * s1[2];
                        if (Q.srctype == 12) {
                            float sphi, cphi;
                            r = 1.f / sqrtf(s1[0] * s1[0] + s1[1] * s1[1] + s1[2] * s1[2]);
                            if (s2[0] > 0.f) {
                                float ax = s1[0] * r, ay = s1[1] * r, az = s1[2] * r;
                                r = v.x * ax + v.y * ay + v.z * az;
                                v.x = v.x - r * ax; v.y = v.y - r * ay; v.z = v.z - r * az;
                                r = 1.f / sqrtf(v.x * v.x + v.y * v.y + v.z * v.z);
                                v.x = v.x * r; v.y = v.y * r; v.z = v.z * r;
                                r = (g_mut == 14) ? s2[0] * t.next() : s2[0] * (2.f * t.next() - 1.f);
                                om_sincos2pi(r * 0.159154943091895336f, &sphi, &cphi);
                                float cx = ay * v.z - az * v.y, cy = az * v.x - ax * v.z, cz = ax * v.y - ay * v.x;
                                v.x = v.x * cphi + cx * sphi;
                                v.y = v.y * cphi + cy * sphi;
                                v.z = v.z * cphi + cz * sphi;
                            } else {
                                v.x = s1[0] * r; v.y = s1[1] * r; v.z = s1[2] * r;
                                om_sincos2pi(t.next(), &sphi, &cphi);
                                rotatevector(v, 1.f, 0.f, sphi, cphi);
                            }
                        } else if (s2[0] > 0.f || s2[1] > 0.f) {
                            float sphi, cphi;
                            om_sincos2pi(t.next(), &sphi, &cphi);
                            r = sqrtf(2.f * t.scatlen());
                            cphi *= s2[g_mut == 15 ? 1 : 0] * r;
                            sphi *= s2[g_mut == 15 ? 0 : 1] * r;
                            sphi *= 1.f / sqrtf(s1[0] * s1[0] + s1[1] * s1[1] + s1[2] * s1[2]);
                            float cx = s1[1] * v.z - s1[2] * v.y, cy = s1[2] * v.x - s1[0] * v.z, cz = s1[0] * v.y - s1[1] * v.x;
                            cphi *= 1.f / sqrtf(cx * cx + cy * cy + cz * cz);
                            v.x += cphi * cx + sphi * s1[0];
                            v.y += cphi * cy + sphi * s1[1];
                            v.z += cphi * cz + sphi * s1[2];
                            r = 1.f / sqrtf(v.x * v.x + v.y * v.y + v.z * v.z);
                            v.x *= r; v.y *= r; v.z *= r;
                        }
                        locate_launch();
                        org[0] = Q.srcpos[0] + s1[0] * 0.5f;
                        org[1] = Q.srcpos[1] + s1[1] * 0.5f;
                        org[2] = Q.srcpos[2] + s1[2] * 0.5f;
                        canfocus = (Q.srctype == 13);
                        break;
                    }
                    case 8:  /* disk */
                    case 17: /* ring */
                    case 3: { /* gaussian */
                        float sphi, cphi, r;
                        if (Q.srctype == 17 && (s1[2] > 0.f || s1[3] > 0.f)) {
                            float phi = fabsf(s1[2] - s1[3]) * t.next() + fminf(s1[2], s1[3]);
                            om_sincos2pi(phi * 0.159154943091895336f, &sphi, &cphi);
                        } else {
                            om_sincos2pi(t.next(), &sphi, &cphi);
                        }
                        if (Q.srctype == 17 && g_mut == 19) {
                            r = t.next() * fabsf(s1[0] - s1[1]) + s1[1];
                        } else if (Q.srctype == 8 || Q.srctype == 17) {
                            r = sqrtf(t.next() * fabsf(s1[0] * s1[0] - s1[1] * s1[1]) + s1[1] * s1[1]);
                        } else if (fabsf(Q.srcdir[3]) < 1e-5f || fabsf(s1[1]) < 1e-5f) {
                            r = sqrtf(0.5f * t.scatlen()) * s1[0];
                        } else {
                            r = s1[0] * s1[0] * 3.14159265358979f / s1[1];
                            r = sqrtf(0.5f * t.scatlen() * (1.f + (Q.srcdir[3] * Q.srcdir[3] / (r * r)))) * s1[0];
                        }
                        if (v.z > -1.f + EPSF && v.z < 1.f - EPSF) {
                            float tmp0 = 1.f - v.z * v.z;
                            float tmp1 = r / sqrtf(tmp0);
                            p.x = p.x + tmp1 * (v.x * v.z * cphi - v.y * sphi);
                            p.y = p.y + tmp1 * (v.y * v.z * cphi + v.x * sphi);
                            p.z = p.z - tmp1 * tmp0 * cphi;
                        } else {
                            p.x += r * cphi;
                            p.y += r * sphi;
                        }
                        locate_launch();
                        break;
                    }
                    case 1:  /* isotropic */
                    case 2:  /* cone */
                    case 7: { /* arcsine */
                        float sphi, cphi, stheta, ctheta;
                        om_sincos2pi(t.next(), &sphi, &cphi);
                        if (Q.srctype == 2) {
                            if (s1[1] > 0.f) {
                                float ang = t.next() * s1[0];
                                om_sincos2pi(ang * 0.159154943091895336f, &stheta, &ctheta);
                            } else {
                                float cmax;
                                float dummy;
                                om_sincos2pi(s1[0] * 0.159154943091895336f, &dummy, &cmax);
                                ctheta = t.next() * (1.f - cmax) + cmax;
                                stheta = om_sin_from_cos(ctheta);
                            }
                        } else if (Q.srctype == 1) {
                            ctheta = 2.f * t.next() - 1.f;
                            stheta = om_sin_from_cos(ctheta);
                        } else {
                            om_sincos2pi(t.next() * (g_mut == 18 ? 0.25f : 0.5f), &stheta, &ctheta);
                        }
                        rotatevector(v, stheta, ctheta, sphi, cphi);
                        canfocus = 0;
                        break;
                    }
                    default:
                        break;
                }
                if (fabsf(p.w) <= Q.minenergy) continue;
                if (canfocus) {
                    if (std::isnan(Q.srcdir[3])) {
                        float sphi, cphi;
                        om_sincos2pi(t.next(), &sphi, &cphi);
                        float ctheta = 2.f * t.next() - 1.f;
                        rotatevector(v, om_sin_from_cos(ctheta), ctheta, sphi, cphi);
                    } else if (Q.srcdir[3] < 0.f && std::isinf(Q.srcdir[3])) {
                        float sphi, cphi;
                        om_sincos2pi(t.next(), &sphi, &cphi);
                        float stheta = sqrtf(t.next());
                        float ctheta = sqrtf(1.f - stheta * stheta);
                        rotatevector(v, stheta, ctheta, sphi, cphi);
                    } else if (Q.srcdir[3] != 0.f) {
                        float Rn2 = (float)((Q.srcdir[3] > 0.f) - (Q.srcdir[3] < 0.f));
                        org[0] += Q.srcdir[3] * v.x;
                        org[1] += Q.srcdir[3] * v.y;
                        org[2] += Q.srcdir[3] * v.z;
                        v.x = Rn2 * (org[0] - p.x);
                        v.y = Rn2 * (org[1] - p.y);
                        v.z = Rn2 * (org[2] - p.z);
                        Rn2 = 1.f / sqrtf(v.x * v.x + v.y * v.y + v.z * v.z);
                        v.x *= Rn2; v.y *= Rn2; v.z *= Rn2;
                    }
                }
            }
            rv[0] = 1.f / v.x; rv[1] = 1.f / v.y; rv[2] = 1.f / v.z;
            if ((mediaid & MED_MASK) == 0) {
                int i1 = skipvoid();
                if (i1 >= 0) {
                    idx1d = (unsigned)i1;
                    mediaid = Q.media[idx1d];
                }
            }
            flipdir[0] = (int)floorf(p.x); flipdir[1] = (int)floorf(p.y); flipdir[2] = (int)floorf(p.z);
            w0 += 1.f;
            if (w0 > (float)Q.maxvoidstep) return -1;
        } while ((mediaid & MED_MASK) == 0 || fabsf(p.w) <= Q.minenergy);
        f.w += 1.f;
        setprop(mediaid, idx1d);
        pp[1] += p.w;
        w0 = p.w;
        pp[2] = p.w;
        v.w = EPSF;
        f.z = 0.f;
        nlaunch++;
        pstep = 0;
        return 0;
    }

    void deposit(unsigned int vox, int tshift, float weight) {
        size_t nvox = (size_t)C->dimlen_z;
        size_t i = (size_t)vox + (size_t)tshift * nvox;
        ndep++;
        if (C->O->field) atomic_add_double(C->O->field + i, (double)weight);
        if (C->O->field32) {
            float* fld = C->O->field32;
            float oldval = atomic_add_float_ret(fld + i, weight);
            if (fabsf(oldval) > MAX_ACCUM) {
                atomic_add_float_ret(fld + i, (oldval > 0.f) ? -MAX_ACCUM : MAX_ACCUM);
                atomic_add_float_ret(fld + i + C->shadowoff,
                                     (oldval > 0.f) ? MAX_ACCUM : -MAX_ACCUM);
            }
        }
    }

    /* src/mcx_core.cu:1767-2528 for one virtual thread */
    void run() {
        const OracleParams& Q = P();
        const Ctx& c = *C;
        const unsigned int flag = Q.savedetflag;
        const unsigned int NSCAT = (flag >> 1) & 1, PPATH = (flag >> 2) & 1, MOM = (flag >> 3) & 1;
        p = {0.f, 0.f, 0.f, NAN};
        v = {0.f, 0.f, 0.f, 0.f};
        f = {0.f, 0.f, 0.f, -1.f};
        flipdir[0] = flipdir[1] = flipdir[2] = 0; flipdir[3] = -1;
        memcpy(&mediaid, &Q.srcparam2[3], 4);
        idx1d = 0;
        unsigned int idx1dold, mediaidold = 0, isdet = 0;
        float n1, len, slen;
        prop = c.gprop[1];
        t.seed(Q.seeds + (size_t)idx * 4);
        if (launch(0)) return;
        rv[0] = 1.f / v.x; rv[1] = 1.f / v.y; rv[2] = 1.f / v.z;
        isdet = mediaid & DET_MASK;
        mediaid &= MED_MASK;
        while (f.w < (float)nphoton) {
            nstep++;
            if (++pstep > MAX_PHOTON_STEPS) {
                err++;
                return;
            }
            if (f.x <= 0.f) {
                f.x = t.scatlen();
                if (v.w != EPSF) {
                    float cphi = 1.f, sphi = 0.f, stheta, ctheta, tmp0;
                    if (!Q.is2d) om_sincos2pi(t.next(), &sphi, &cphi);
                    tmp0 = (v.w > (float)Q.gscatter) ? 0.f : prop.z;
                    ctheta = scatter_costheta(tmp0, prop.z, t.next());
                    stheta = om_sin_from_cos(ctheta);
                    if (Q.savedet) {
                        const unsigned int lab = split() ? sv_label() : (mediaid & MED_MASK);
                        if (!split() || lab > 0) {
                            if (NSCAT) ppath[lab - 1] += 1.f;
                            if (MOM) ppath[c.maxmedia * (NSCAT + PPATH) + lab - 1] += 1.f - ctheta;
                        }
                    }
                    if (Q.is2d)
                        rotatevector2d(v, Q.is2d, (t.next() > 0.5f ? stheta : -stheta), ctheta);
                    else
                        rotatevector(v, stheta, ctheta, sphi, cphi);
                    v.w += 1.f;
                    rv[0] = 1.f / v.x; rv[1] = 1.f / v.y; rv[2] = 1.f / v.z;
                    nscat++;
                    /* src/mcx_core.cu:2019-2062 */
                    if (Q.replayseed && (Q.outputtype == otWP || Q.outputtype == otDCS || Q.outputtype == otWPTOF ||
                                         Q.outputtype == otRFmus)) {
                        const float tof = Q.replaytof[gid];
                        int ts = (int)floorf((tof - Q.twin0) * Q.Rtstep);
                        ts = std::max(0, std::min((int)Q.maxgate - 1, ts));
                        if (Q.replaydet == -1) ts += ((Q.replaydetid[gid] & 0xFFFF) - 1) * (int)Q.maxgate;
                        if (Q.outputtype == otRFmus) {
                            /* the detected weight times the modulation phasor
                             * of the photon's time of flight */
                            float rs, rc;
                            rf_sincos(Q.omega * tof, &rs, &rc);
                            if (g_mut == 20) std::swap(rs, rc);
                            const float rw = Q.replayweight[gid];
                            deposit(idx1d, ts, rw * rc);
                            if (Q.omega > 0.f) deposit(idx1d, ts + c.rfshift, rw * rs);
                        } else {
                            float w = (Q.outputtype == otDCS) ? (1.f - ctheta) : 1.f;
                            w = (Q.outputtype == otWPTOF) ? tof : w;
                            w *= Q.replayweight[gid];
                            deposit(idx1d, ts, w);
                        }
                    }
                }
                v.w = (float)(int)v.w;
                testint = true;
            }
            n1 = prop.w;
            if (!split() || !svsplit) setprop(mediaid, idx1d);
            len = hitgrid(p, v, rv, flipdir);
            const bool gs = (v.w + 1.f > (float)Q.gscatter);
            slen = len * prop.y * (gs ? (1.f - prop.z) : 1.f);
            slen = fminf(slen, f.x);
            len = ((prop.y == 0.f) ? len : (slen / (prop.y * (gs ? (1.f - prop.z) : 1.f))));
            /* src/mcx_core.cu:2112-2117: the interface plane of a split voxel */
            hitintf = split() && svsplit && testint && ray_plane(len, slen);
            advance(p, v, len, (slen == f.x) || hitintf, flipdir);
            p.w *= om_expf(-prop.x * len);
            f.x -= slen;
            f.y += len * prop.w * Q.oneoverc0;
            f.z += len;
            if (Q.savedet && PPATH) {
                const unsigned int lab = split() ? sv_label() : (mediaid & MED_MASK);
                if (!split() || lab > 0) ppath[c.maxmedia * NSCAT + lab - 1] += len;
            }
            mediaidold = mediaid | (isdet & DET_MASK);
            idx1dold = idx1d;
            idx1d = (unsigned)(flipdir[2] * (int)c.dimlen_y + flipdir[1] * (int)c.dimlen_x + flipdir[0]);
            if ((unsigned)flipdir[0] >= Q.nx || (unsigned)flipdir[1] >= Q.ny || (unsigned)flipdir[2] >= Q.nz) {
                mediaid = 0;
                idx1d = (flipdir[0] < 0 || flipdir[1] < 0 || flipdir[2] < 0) ? OUT_MIN : OUT_MAX;
                isdet = Q.bc[(idx1d == OUT_MAX) * 3 + flipdir[3]];
            } else {
                mediaid = Q.media[idx1d];
                isdet = mediaid & DET_MASK;
                mediaid &= MED_MASK;
            }
            if ((idx1d != idx1dold || hitintf) && mediaidold) {
                if (Q.save2pt && f.y >= Q.twin0 && f.y < Q.twin1) {
                    float weight = 0.f, imag = 0.f;
                    int tshift = (int)floorf((f.y - Q.twin0) * Q.Rtstep);
                    if (tshift > (int)Q.maxgate - 1) tshift = (int)Q.maxgate - 1;
                    if (Q.outputtype == otEnergy)
                        weight = w0 - p.w;
                    else if (Q.outputtype == otFluence || Q.outputtype == otFlux)
                        weight = (prop.x < EPSF) ? (w0 * f.z) : ((w0 - p.w) / prop.x);
                    else if (Q.replayseed && (Q.outputtype == otJacobian || Q.outputtype == otRF || Q.outputtype == otWLTOF)) {
                        /* src/mcx_core.cu:2208-2220, gate clamped to the range */
                        const float tof = Q.replaytof[gid];
                        weight = Q.replayweight[gid] * f.z;
                        if (Q.outputtype == otRF) {
                            /* src/mcx_core.cu:2211-2213, :2254-2257, :2271-2273:
                             * real part -w L cos(omega tof); imaginary part
                             * -w L sin(omega tof), or p.w L sin(omega tof)
                             * for pattern sources */
                            float rs, rc;
                            rf_sincos(Q.omega * tof, &rs, &rc);
                            if (g_mut == 20) std::swap(rs, rc);
                            imag = (Q.srctype == 5 || Q.srctype == 15) ? p.w * f.z * rs : -weight * rs;
                            weight = -weight * rc;
                        } else if (Q.outputtype == otWLTOF) {
                            weight = weight * tof;
                        }
                        tshift = std::max(0, std::min((int)Q.maxgate - 1, (int)floorf((tof - Q.twin0) * Q.Rtstep)));
                        if (Q.replaydet == -1) tshift += ((Q.replaydetid[gid] & 0xFFFF) - 1) * (int)Q.maxgate;
                    } else if (Q.outputtype == otL)
                        weight = w0 * f.z;
                    if (fabsf(weight) > 0.f) deposit(idx1dold, tshift, weight);
                    if (c.rfshift && Q.omega > 0.f && fabsf(imag) > 0.f) deposit(idx1dold, tshift + c.rfshift, imag);
                }
                w0 = p.w;
                f.z = 0.f;
            } else {
                mediaid = mediaidold;
            }
            /* src/mcx_core.cu:2294-2304 */
            if (split()) {
                if (idx1d != idx1dold) {
                    setprop(mediaid, idx1d);
                    testint = true;
                } else if (hitintf) {
                    for (int k = 0; k < 3; k++) nv[k] = -nv[k];
                    svupper = !svupper;
                    testint = false;
                }
            }
            const unsigned int bcf = isdet & 0xF;
            if ((mediaid == 0 &&
                 ((!Q.isreflect || (Q.isreflect && n1 == c.gprop[0].w)) || (bcf == bcUnknown && !Q.doreflect) ||
                  bcf == bcAbsorb || bcf == bcCyclic) &&
                 bcf != bcMirror) ||
                (split() && (idx1d != idx1dold || hitintf) && !svupper && !svlo &&
                 (!Q.isreflect || (Q.isreflect && n1 == c.gprop[0].w))) ||
                f.y > Q.twin1) {
                if (isdet == bcCyclic) {
                    if (flipdir[3] == 0) {
                        p.x = nextafter_mcx(roundf(p.x + ((idx1d == OUT_MIN) ? c.maxidx_x : -c.maxidx_x)), (v.x > 0.f) - (v.x < 0.f), Q.maxvoidstep);
                        flipdir[0] = (int)floorf(p.x);
                    }
                    if (flipdir[3] == 1) {
                        p.y = nextafter_mcx(roundf(p.y + ((idx1d == OUT_MIN) ? c.maxidx_y : -c.maxidx_y)), (v.y > 0.f) - (v.y < 0.f), Q.maxvoidstep);
                        flipdir[1] = (int)floorf(p.y);
                    }
                    if (flipdir[3] == 2) {
                        p.z = nextafter_mcx(roundf(p.z + ((idx1d == OUT_MIN) ? c.maxidx_z : -c.maxidx_z)), (v.z > 0.f) - (v.z < 0.f), Q.maxvoidstep);
                        flipdir[2] = (int)floorf(p.z);
                    }
                    if ((unsigned)flipdir[0] < Q.nx && (unsigned)flipdir[1] < Q.ny && (unsigned)flipdir[2] < Q.nz) {
                        idx1d = (unsigned)(flipdir[2] * (int)c.dimlen_y + flipdir[1] * (int)c.dimlen_x + flipdir[0]);
                        mediaid = Q.media[idx1d];
                        isdet = mediaid & DET_MASK;
                        mediaid &= MED_MASK;
                        continue;
                    }
                }
                unsigned int detarg = (((idx1d == OUT_MAX && Q.bc[9 + flipdir[3]]) || (idx1d == OUT_MIN && Q.bc[6 + flipdir[3]]))
                                           ? OUT_MIN : (mediaidold & DET_MASK));
                if (launch(detarg)) break;
                isdet = mediaid & DET_MASK;
                mediaid &= MED_MASK;
                continue;
            }
            if (fabsf(p.w) < Q.minenergy) {
                if (t.next() * ROULETTE <= 1.f) {
                    p.w *= ROULETTE;
                } else {
                    if (launch(mediaidold & DET_MASK)) break;
                    isdet = mediaid & DET_MASK;
                    mediaid &= MED_MASK;
                    continue;
                }
            }
            if (Q.isreflect && hitintf) {
                /* src/mcx_core.cu:2386-2404 */
                if (c.gprop[svlo].w != c.gprop[svup].w) {
                    for (int k = 0; k < 3; k++) nv[k] = -nv[k];
                    if (reflectray(n1)) {
                        if (launch(mediaidold & DET_MASK)) break;
                        isdet = mediaid & DET_MASK;
                        mediaid &= MED_MASK;
                        continue;
                    }
                } else {
                    prop = c.gprop[sv_label()];
                }
            } else if (Q.isreflect) {
                /* src/mcx_core.cu:2381-2416 */
                const bool lt100 = Q.mediaformat < 100;
                if (lt100 && !split()) setprop(mediaid, idx1d);
                if (((mediaid && Q.doreflect) ||
                     (mediaid == 0 && ((bcf == bcUnknown && Q.doreflect) || bcf == bcReflect || bcf == bcMirror))) &&
                    (bcf == bcMirror || n1 != (lt100 ? prop.w : c.gprop[mediaid > 0u ? 1 : 0].w))) {
                    if (!split()) setprop(mediaid, idx1d);
                    const float Rtotal = (bcf == bcMirror) ? 1.f : fresnel_face(n1, prop.w, fabsf(comp(v, flipdir[3])));
                    if (Rtotal < 1.f && (!(mediaid == 0 && bcf == bcMirror)) && t.next() > Rtotal) {
                        transmit(v, n1, prop.w, flipdir[3]);
                        if (mediaid == 0 || (split() && sv_label() == 0)) {
                            unsigned int detarg = (((idx1d == OUT_MAX && Q.bc[9 + flipdir[3]]) || (idx1d == OUT_MIN && Q.bc[6 + flipdir[3]]))
                                                       ? OUT_MIN : (mediaidold & DET_MASK));
                            if (launch(detarg)) break;
                            isdet = mediaid & DET_MASK;
                            mediaid &= MED_MASK;
                            continue;
                        }
                        rv[0] = 1.f / v.x; rv[1] = 1.f / v.y; rv[2] = 1.f / v.z;
                    } else {
                        int fd = flipdir[3];
                        if (fd == 0) v.x = -v.x; else if (fd == 1) v.y = -v.y; else v.z = -v.z;
                        rv[0] = 1.f / v.x; rv[1] = 1.f / v.y; rv[2] = 1.f / v.z;
                        if (fd == 0) {
                            p.x = nextafter_mcx(rintf(p.x), (v.x > 0.f) - (v.x < 0.f), Q.maxvoidstep);
                            flipdir[0] = (int)floorf(p.x);
                        } else if (fd == 1) {
                            p.y = nextafter_mcx(rintf(p.y), (v.y > 0.f) - (v.y < 0.f), Q.maxvoidstep);
                            flipdir[1] = (int)floorf(p.y);
                        } else {
                            p.z = nextafter_mcx(rintf(p.z), (v.z > 0.f) - (v.z < 0.f), Q.maxvoidstep);
                            flipdir[2] = (int)floorf(p.z);
                        }
                        idx1d = idx1dold;
                        mediaid = Q.media[idx1d] & MED_MASK;
                        setprop(mediaid, idx1d);
                        if (split() && sv_label() == 0) {
                            if (launch(mediaidold & DET_MASK)) break;
                            isdet = mediaid & DET_MASK;
                            mediaid &= MED_MASK;
                            continue;
                        }
                        n1 = prop.w;
                    }
                } else if (lt100 && !split()) {
                    setprop(mediaidold, idx1d);
                }
            } else if (split()) {
                prop = c.gprop[sv_label()];
            }
            if (mediaid == 0 || idx1d == OUT_MIN || idx1d == OUT_MAX) {
                err++;
                return;
            }
        }
        C->O->genergy[idx * 2] = ppath[c.partialdata];
        C->O->genergy[idx * 2 + 1] = ppath[c.partialdata + 1];
    }
};

}  // namespace

extern "C" {

float mcx_oracle_expf(float x) { return om_expf(x); }
float mcx_oracle_logf(float x) { return om_logf(x); }
void mcx_oracle_sincos2pi(float u, float* s, float* c) { om_sincos2pi(u, s, c); }
void mcx_oracle_sincosf(float x, float* s, float* c) { rf_sincos(x, s, c); }

void mcx_oracle_glibc_rand(unsigned int seed, unsigned int n, int* out) {
    GlibcRand g(seed);
    for (unsigned int i = 0; i < n; i++) out[i] = g.next();
}

void mcx_oracle_seedwords(unsigned int seed, unsigned int nwords, unsigned int* words) {
    GlibcRand g(seed);
    for (unsigned int i = 0; i < nwords; i++) {
        uint32_t a = (uint32_t)g.next();
        uint32_t b = (uint32_t)g.next();
        uint32_t c = (uint32_t)g.next();
        words[i] = (a << 16) | (b << 1) | (c >> 14);
    }
}

void mcx_oracle_rng_floats(const unsigned int* seed4, unsigned int n, float* out) {
    Rng r;
    r.seed(seed4);
    for (unsigned int i = 0; i < n; i++) out[i] = r.next();
}

void mcx_oracle_maskdet(unsigned int* vol, unsigned int nx, unsigned int ny, unsigned int nz,
                        const float* detpos, unsigned int detnum, int svmc) {
    const float corners[8][3] = {{0, 0, 0}, {1, 0, 0}, {0, 1, 0}, {0, 0, 1}, {1, 1, 0}, {1, 0, 1}, {0, 1, 1}, {1, 1, 1}};
    unsigned int dx = nx + 2, dy = ny + 2, dz = nz + 2;
    int isonecube = (nx == 1) && (ny == 1) && (nz == 1);
    std::vector<unsigned int> pad((size_t)dx * dy * dz, 0);
    for (unsigned int z = 1; z <= nz; z++)
        for (unsigned int y = 1; y <= ny; y++)
            memcpy(&pad[(size_t)z * dy * dx + y * dx + 1], vol + (size_t)(z - 1) * ny * nx + (y - 1) * nx, nx * 4);
    for (unsigned int d = 0; d < detnum; d++) {
        const float* dp = detpos + 4 * d;
        float radius = fabsf(dp[3]);
        float d2max = (radius + 1.7321f) * (radius + 1.7321f);
        for (float z = -radius - 1.f; z <= radius + 1.f; z += 0.5f) {
            float iz = z + dp[2];
            for (float y = -radius - 1.f; y <= radius + 1.f; y += 0.5f) {
                float iy = y + dp[1];
                for (float x = -radius - 1.f; x <= radius + 1.f; x += 0.5f) {
                    float ix = x + dp[0];
                    if (iz < 0 || ix < 0 || iy < 0 || ix >= nx || iy >= ny || iz >= nz ||
                        x * x + y * y + z * z > (radius + 1.f) * (radius + 1.f))
                        continue;
                    float mind2 = 1.f / EPSF;
                    for (int c = 0; c < 8; c++) {
                        float rx = (int)ix - dp[0] + corners[c][0];
                        float ry = (int)iy - dp[1] + corners[c][1];
                        float rz = (int)iz - dp[2] + corners[c][2];
                        float d2 = rx * rx + ry * ry + rz * rz;
                        if (d2 > d2max) { mind2 = 1.f / EPSF; break; }
                        if (d2 < mind2) mind2 = d2;
                    }
                    float lim = radius + 0.5f * (1.f + isonecube);
                    if (mind2 == 1.f / EPSF || mind2 >= lim * lim) continue;
                    size_t i = (size_t)(int)(iz + 1.f) * dy * dx + (size_t)(int)(iy + 1.f) * dx + (size_t)(int)(ix + 1.f);
                    /* split voxels (src/mcx_utils.c:4287-4310): background is a
                     * lower label of 0 */
                    const unsigned int lower = (pad[i] & 0xFF000000u) >> 24, upper = (pad[i] & 0x00FF0000u) >> 16;
                    if (svmc ? !(lower || upper) : !pad[i]) continue;
                    bool interior = !(svmc && !lower && upper);
                    for (int oz = -1; oz <= 1 && interior; oz++)
                        for (int oy = -1; oy <= 1 && interior; oy++)
                            for (int ox = -1; ox <= 1; ox++) {
                                int nzero = (ox != 0) + (oy != 0) + (oz != 0);
                                if (nzero == 0) continue;
                                size_t j = i + (long)oz * dy * dx + (long)oy * dx + ox;
                                if (svmc ? !((pad[j] & 0xFF000000u) >> 24) : !pad[j]) { interior = false; break; }
                            }
                    if (!interior)
                        vol[(size_t)(int)iz * ny * nx + (size_t)(int)iy * nx + (int)ix] |= DET_MASK;
                }
            }
        }
    }
}

}  // extern "C"

namespace {

/* The launch constants of src/mcx_core.cu:2853-2876 and the property +
 * detector table (:3316-3321) for one run. */
int init_ctx(const OracleParams* P, OracleOut* O, Ctx& c, std::vector<Vec4>& gprop) {
    if (!P || !O || !P->media || !P->prop || !P->seeds) return -1;
    if (P->medianum < 2 || P->medianum + P->detnum > 4000) return -2;
    c.P = P;
    c.O = O;
    c.dimlen_x = P->nx;
    c.dimlen_y = P->nx * P->ny;
    c.dimlen_z = P->nx * P->ny * P->nz;
    c.maxidx_x = (float)P->nx; c.maxidx_y = (float)P->ny; c.maxidx_z = (float)P->nz;
    c.maxmedia = P->medianum - 1;
    unsigned int f = P->savedetflag;
    c.partialdata = c.maxmedia * (((f >> 1) & 1) + ((f >> 2) & 1) + ((f >> 3) & 1));
    c.w0offset = c.partialdata + 4;
    c.reclen = c.partialdata + (f & 1) + 3 * (((f >> 4) & 1) + ((f >> 5) & 1)) + ((f >> 6) & 1);
    O->reclen = c.reclen;
    const unsigned int blocks = (P->replayseed && P->replaydet == -1) ? P->detnum : 1;
    const bool rf = P->replayseed && (P->outputtype == otRF || P->outputtype == otRFmus);
    c.rfshift = rf ? (int)(P->maxgate * blocks) : 0;
    c.shadowoff = (size_t)c.dimlen_z * P->maxgate * blocks * (rf ? 2 : 1);
    /* the reference's numbering reads the photon before the first when no
     * thread takes an extra photon */
    if (P->replayseed && P->replayindex && P->oddphotons == 0 && P->threadphoton > 0) return -3;
    gprop.assign(P->medianum + P->detnum, Vec4{0.f, 0.f, 0.f, 0.f});
    for (unsigned int i = 0; i < P->medianum; i++) gprop[i] = {P->prop[4 * i], P->prop[4 * i + 1], P->prop[4 * i + 2], P->prop[4 * i + 3]};
    for (unsigned int i = 0; i < P->detnum; i++)
        gprop[P->medianum + i] = {P->detpos[4 * i], P->detpos[4 * i + 1], P->detpos[4 * i + 2], P->detpos[4 * i + 3]};
    c.gprop = gprop.data();
    return 0;
}

}  // namespace

extern "C" {

int mcx_oracle_run(const OracleParams* P, OracleOut* O) {
    if (!O || !O->genergy) return -1;
    Ctx c;
    std::vector<Vec4> gprop;
    const int rc = init_ctx(P, O, c, gprop);
    if (rc) return rc;
    std::atomic<unsigned int> detected(O->detected);
    c.detected = &detected;
    unsigned long long nstep = 0, ndep = 0, nscat = 0, nlaunch = 0;
    int errors = 0;
    O->errors = 0;
    const unsigned int pplen = c.w0offset + 1;
    long tb = P->thread_begin, te = P->thread_end ? P->thread_end : P->nthread;
    if (te > (long)P->nthread) te = P->nthread;
    const long total_needed = (long)P->threadphoton * (long)P->nthread + P->oddphotons;
    auto t0 = std::chrono::steady_clock::now();
#ifdef _OPENMP
    int nt = P->ompthreads > 0 ? P->ompthreads : omp_get_max_threads();
#pragma omp parallel for schedule(dynamic, 16) num_threads(nt) reduction(+ : nstep, ndep, nscat, nlaunch, errors)
#endif
    for (long i = tb; i < te; i++) {
        if (i >= total_needed) continue;
        std::vector<float> pp(pplen, 0.f);
        Thread T;
        T.C = &c;
        T.idx = (unsigned)i;
        T.nphoton = P->threadphoton + (i < P->oddphotons);
        T.ppath = pp.data();
        T.ppath[c.partialdata] = O->genergy[2 * i];
        T.ppath[c.partialdata + 1] = O->genergy[2 * i + 1];
        T.run();
        nstep += T.nstep; ndep += T.ndep; nscat += T.nscat; nlaunch += T.nlaunch;
        if (O->thread_steps) O->thread_steps[i] = T.nstep;
        errors += T.err;
    }
    auto t1 = std::chrono::steady_clock::now();
    O->seconds = std::chrono::duration<double>(t1 - t0).count();
    O->detected = detected.load();
    O->nstep = nstep; O->ndeposit = ndep; O->nscatter = nscat; O->nlaunch = nlaunch;
    O->errors += errors;
    return 0;
}

void mcx_oracle_set_mutation(int id) { g_mut = id; }

float mcx_oracle_hg_costheta(float g, float u) { return scatter_costheta(g, g, u); }

float mcx_oracle_fresnel_face(float n1, float n2, float cosi) { return fresnel_face(n1, n2, cosi); }

float mcx_oracle_reflectcoeff(float n1, float n2, float cosi) {
    const Vec4 v = {0.f, 0.f, cosi, 0.f};
    return reflectcoeff(v, n1, n2, 2);
}

void mcx_oracle_transmit(float n1, float n2, int face, const float* vin, float* vout) {
    Vec4 v = {vin[0], vin[1], vin[2], 0.f};
    transmit(v, n1, n2, face);
    vout[0] = v.x; vout[1] = v.y; vout[2] = v.z;
}

int mcx_oracle_trace(const float* p0, const float* v0, unsigned int nx, unsigned int ny, unsigned int nz,
                     unsigned int maxn, int* ids, float* lens) {
    Vec4 p = {p0[0], p0[1], p0[2], 1.f};
    const Vec4 v = {v0[0], v0[1], v0[2], 0.f};
    const float rv[3] = {1.f / v.x, 1.f / v.y, 1.f / v.z};
    int fd[4] = {(int)floorf(p.x), (int)floorf(p.y), (int)floorf(p.z), -1};
    unsigned int n = 0;
    while (n < maxn) {
        const float len = hitgrid(p, v, rv, fd);
        advance(p, v, len, false, fd);
        ids[3 * n] = fd[0]; ids[3 * n + 1] = fd[1]; ids[3 * n + 2] = fd[2];
        lens[n] = len;
        n++;
        if ((unsigned)fd[0] >= nx || (unsigned)fd[1] >= ny || (unsigned)fd[2] >= nz) break;
    }
    return (int)n;
}

int mcx_oracle_launch_samples(const OracleParams* P, unsigned int n, float* out) {
    OracleOut O;
    memset(&O, 0, sizeof(O));
    Ctx c;
    std::vector<Vec4> gprop;
    const int rc = init_ctx(P, &O, c, gprop);
    if (rc) return rc;
    std::atomic<unsigned int> detected(0);
    c.detected = &detected;
    std::vector<float> pp(c.w0offset + 1, 0.f);
    Thread T;
    T.C = &c;
    T.idx = 0;
    T.nphoton = INT_MAX;
    T.ppath = pp.data();
    T.p = {0.f, 0.f, 0.f, NAN};
    T.v = {0.f, 0.f, 0.f, 0.f};
    T.f = {0.f, 0.f, 0.f, -1.f};
    T.flipdir[0] = T.flipdir[1] = T.flipdir[2] = 0;
    T.flipdir[3] = -1;
    memcpy(&T.mediaid, &P->srcparam2[3], 4);
    T.idx1d = 0;
    T.prop = c.gprop[1];
    T.t.seed(P->seeds);
    for (unsigned int i = 0; i < n; i++) {
        if (T.launch(0)) return -3;
        float* o = out + 8 * (size_t)i;
        o[0] = T.p.x; o[1] = T.p.y; o[2] = T.p.z; o[3] = T.p.w;
        o[4] = T.v.x; o[5] = T.v.y; o[6] = T.v.z; o[7] = T.v.w;
        T.f.w = 0.f;
    }
    return (int)n;
}

}  // extern "C"
