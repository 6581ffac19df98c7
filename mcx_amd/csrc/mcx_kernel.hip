/*
 * mcx_kernel.hip -- photon-transport kernel for MI355X (gfx950, wave64).
 *
 * One lane owns one xorshift128+ stream and runs its share of photons one
 * after another, exactly as the reference's virtual threads do
 * (src/mcx_core.cu:1767-2528, :3014-3022); the per-step float work follows
 * the reference's order of operations, with the elementary functions of
 * mcx_math.h, so a lane's trajectories are reproducible bit for bit by the
 * CPU restatement used in the tests.
 *
 * MI355X choices:
 *  - all launch constants arrive in the kernel argument block (scalar
 *    registers, uniform branches) instead of constant memory;
 *  - the media/detector property table is copied once per workgroup into
 *    LDS and read with ds_read_b128 by the per-lane medium index;
 *  - the label volume is uint8 (label | detector bit 7) when the labels fit,
 *    4x less traffic than the reference's uint32 volume;
 *  - detector partial paths live in LDS in lane-interleaved (SoA) order,
 *    so a wave's accesses hit 64 consecutive dwords (no bank conflicts);
 *  - fluence deposits use hardware global_atomic_add_f32 (returning only
 *    where the MAX_ACCUM shadow needs the old value) or f64 atomics;
 *  - detected photons reserve record slots with one atomic per wave
 *    (ballot + prefix count) instead of one per lane.
 */
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/mcx_media.h"
#include "mcx_kernel.h"
#include "mcx_math.h"

namespace mcxk {

constexpr float EPSF = 1.1920928955078125e-07f;
constexpr uint32_t DET_MASK = 0x80000000u;
constexpr uint32_t MED_MASK = 0x7FFFFFFFu;
constexpr uint32_t OUT_MIN = 0xFFFFFFFFu;
constexpr uint32_t OUT_MAX = 0x7FFFFFFFu;
constexpr float MAX_ACCUM = 1000.f;
constexpr float ROULETTE = 10.f;
constexpr float JUST_BELOW_ONE = 0.9998f;
/* Steps one photon may take before its lane stops on an internal error
 * (counted like a photon left outside the volume).  With time windows of
 * nanoseconds a history takes at most ~1e5 steps; the limit turns a
 * history that stops advancing into a reported error, not a launch that
 * never ends. */
constexpr uint32_t MAX_PHOTON_STEPS = 1u << 26;
enum { bcUnknown, bcReflect, bcAbsorb, bcMirror, bcCyclic };
enum { otFlux, otFluence, otEnergy, otJacobian, otWP, otDCS, otRF, otL, otRFmus, otWLTOF, otWPTOF };

struct F4 { float x, y, z, w; };

/* The buffers arrive through the KParam block in device memory, so the
 * compiler cannot see that they are global memory and would emit FLAT
 * instructions, whose completion waits on both the vector-memory and the
 * LDS/scalar counters.  Every buffer access goes through these casts to
 * the global address space (global_load / global_atomic instructions). */
#define MCX_GLB __attribute__((address_space(1)))
template <typename T>
__device__ __forceinline__ MCX_GLB T* glb(T* p) { return (MCX_GLB T*)p; }

__device__ __forceinline__ void atom_add_nr(double* p, double v) {
    __hip_atomic_fetch_add(glb(p), v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void atom_add_nr(float* p, float v) {
    __hip_atomic_fetch_add(glb(p), v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ float atom_add_ret(float* p, float v) {
    return __hip_atomic_fetch_add(glb(p), v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ uint32_t atom_add_u32(uint32_t* p, uint32_t v) {
    return __hip_atomic_fetch_add(glb(p), v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void atom_add_nr(int32_t* p, int32_t v) {
    __hip_atomic_fetch_add(glb(p), v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

/* Completes an FP32 hi/lo deposit of w whose returning add found `a` in
 * the hi half of a cell (the add stored s = fl(a + w)):
 *  - the exact rounding error of that add (TwoSum) goes to the lo half;
 *  - when a -> s crossed a multiple of MCX_HILO_CHUNK, the lane moves the
 *    chunks it crossed from the hi half to the cell's chunk count, by a
 *    returning add whose rounding error also goes to the lo half.
 * Every move is claimed by the one add that crossed the boundary, so the
 * hi half stays within about a chunk of zero and its rounding step far
 * below the deposits.  The same lane then takes the lo half (exchange with
 * zero) and adds it to the hi half, again a returning add whose error goes
 * back to lo: the lo half is emptied every chunk, so it stays small even
 * when its errors are one-signed (equal deposits, e.g. every unscattered
 * photon crossing the source voxel, round the same way each time).
 * Without both bounds the hi half grows until deposits fall below half its
 * rounding step, and the lo half, grown large, loses what it gathers.
 * The merged value is chunks x MCX_HILO_CHUNK + hi + lo, in FP64. */
__device__ __forceinline__ void hilo_complete(float* hi, float* lo, int32_t* chunks, float a, float w) {
    const float s = a + w;
    const float bb = s - a;
    const float e = (a - (s - bb)) + (w - bb);
    if (e != 0.f) atom_add_nr(lo, e);
    const float m = floorf(s * (1.f / MCX_HILO_CHUNK)) - floorf(a * (1.f / MCX_HILO_CHUNK));
    if (m != 0.f) {
        const float d = -m * MCX_HILO_CHUNK;
        const float a2 = atom_add_ret(hi, d);
        const float s2 = a2 + d;
        const float b2 = s2 - a2;
        const float e2 = (a2 - (s2 - b2)) + (d - b2);
        if (e2 != 0.f) atom_add_nr(lo, e2);
        atom_add_nr(chunks, (int32_t)m);
        const float x = __hip_atomic_exchange(glb(lo), 0.f, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (x != 0.f) {
            const float a3 = atom_add_ret(hi, x);
            const float s3 = a3 + x;
            const float b3 = s3 - a3;
            const float e3 = (a3 - (s3 - b3)) + (x - b3);
            if (e3 != 0.f) atom_add_nr(lo, e3);
        }
    }
}

template <typename MT>
__device__ __forceinline__ uint32_t load_media(const void* m, uint32_t i);

template <>
__device__ __forceinline__ uint32_t load_media<uint8_t>(const void* m, uint32_t i) {
    uint32_t b = glb((const uint8_t*)m)[i];
    return (b & 0x7Fu) | ((b & 0x80u) << 24);
}

template <>
__device__ __forceinline__ uint32_t load_media<uint32_t>(const void* m, uint32_t i) {
    return glb((const uint32_t*)m)[i];
}

/* media type tag of the non-label formats: 32-bit words decoded by
 * kp.mediaformat (ASGN_F2H keeps a second word per voxel nvox words on) */
struct MediaWord {};

template <>
__device__ __forceinline__ uint32_t load_media<MediaWord>(const void* m, uint32_t i) {
    return glb((const uint32_t*)m)[i];
}

/* media type tag of the split-voxel format: per voxel a word with the two
 * labels and two interface-point bytes, and (nvox words on) a word with
 * the third point byte and the normal bytes */
struct MediaSplit {};

template <>
__device__ __forceinline__ uint32_t load_media<MediaSplit>(const void* m, uint32_t i) {
    return glb((const uint32_t*)m)[i];
}

template <typename T> struct is_word_media { static constexpr bool value = false; };
template <> struct is_word_media<MediaWord> { static constexpr bool value = true; };
template <typename T> struct is_split_media { static constexpr bool value = false; };
template <> struct is_split_media<MediaSplit> { static constexpr bool value = true; };

/* IEEE binary16 bits -> float (exact) */
__device__ __forceinline__ float half_to_float(uint32_t h) {
    return (float)__builtin_bit_cast(_Float16, (uint16_t)h);
}

/* XCD executing this wave: s_getreg_b32 hwreg(HW_REG_XCC_ID, 0, 4),
 * masked so the copy index is always in range */
__device__ __forceinline__ uint32_t xcc_id() {
    return (uint32_t)__builtin_amdgcn_s_getreg((3 << 11) | (0 << 6) | 20) & (MCX_NUM_XCD - 1);
}

__device__ __forceinline__ float sgnf(float x) { return (float)((x > 0.f) - (x < 0.f)); }

__device__ __forceinline__ float nextafter_mcx(float a, int dir, uint32_t maxvoidstep) {
    float f = a + (float)maxvoidstep;
    uint32_t i = mcx_f2bits(f);
    i += (uint32_t)dir ^ (i & 0x80000000u);
    return mcx_bits2f(i) - (float)maxvoidstep;
}

/* element of the blocked FP64 tally (KParam::tblock) for voxel (x, y, z)
 * of gate plane g */
__device__ __forceinline__ size_t block_cell(const KParam& kp, uint32_t x, uint32_t y, uint32_t z, uint32_t g) {
    const uint32_t b = ((g * kp.bnz + (z >> 1)) * kp.bny + (y >> 1)) * kp.bnx + (x >> 1);
    return ((size_t)b << 3) + (x & 1u) + ((y & 1u) << 1) + ((z & 1u) << 2);
}

/* voxel index -> x, y, z by the multiply-high divisions of KParam */
__device__ __forceinline__ void vox_xyz(const KParam& kp, uint32_t vox, uint32_t& x, uint32_t& y, uint32_t& z) {
    const uint32_t tz = __umulhi(vox, kp.hot_mxy);
    z = (tz + ((vox - tz) >> (kp.hot_shxy & 0xffu))) >> (kp.hot_shxy >> 8);
    const uint32_t r = vox - z * kp.dimxy;
    const uint32_t ty = __umulhi(r, kp.hot_mx);
    y = (ty + ((r - ty) >> (kp.hot_shx & 0xffu))) >> (kp.hot_shx >> 8);
    x = r - y * kp.nx;
}

/* number of set bits of m below this lane */
__device__ __forceinline__ uint32_t lanes_below(uint64_t m) {
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}

#define MCX_NO_BLOCK 0xFFFFFFFFu

template <int ISPENCIL, int ISREFLECT, int TALLY, typename MT>
struct Lane {
    const KParam& kp;
    const F4* tab;        /* LDS: properties then detectors */
    float* pp;            /* LDS: this lane's partial-path column */
    uint32_t stride;      /* blockDim.x */
    uint32_t idx;
    int nphoton;

    F4 p, v, f;           /* f = {pscat, t, pathlen, ndone} */
    float rv0, rv1, rv2;
    F4 prop;
    uint64_t t0, t1;
    int fd0, fd1, fd2, fd3;
    uint32_t idx1d, mediaid;
    float w0;
    float esc, launched, winit;
    /* FP32 tallies: the returning atomic of the previous deposit, consumed
     * at the next deposit (or lane exit) so its latency overlaps a step */
    size_t pend_i;
    float pend_old, pend_w;
    double* tally64;      /* FP64 modes: this lane's tally copy */
    uint32_t gid;         /* global index of the current photon (replay) */
    uint32_t nstep, ndep, nscat, nlaunch, nhot;
    uint32_t nreq;        /* global tally atomic requests (a block flush is one) */
    uint32_t pstep;       /* steps of the current photon */
    uint32_t hoff;        /* byte offset of this lane's hot-cell box in the dynamic LDS */
    /* blocked FP64 tally (single-simulation kernels): this lane's LDS block
     * (8 cells + block index), the index of the block it holds, the
     * workgroup's block area and this wave's lane queue */
    static constexpr bool BLOCKED = TALLY == MCX_TALLY_F64 && ISPENCIL != 2;
    float* blk;
    uint32_t bopen;
    /* split-voxel state (src/mcx_core.h:65-83): interface point and
     * normal of the current voxel, its two labels, which side the photon
     * is on, whether the voxel is split; whether the current path segment
     * still needs the interface test and whether this step hit it */
    float rpx, rpy, rpz, nvx, nvy, nvz;
    uint32_t svlo, svup, svupper, svsplit;
    bool testint, hitintf;

    __device__ __forceinline__ Lane(const KParam& k, const F4* tb, float* ppb, uint32_t s, uint32_t id)
        : kp(k), tab(tb), pp(ppb), stride(s), idx(id), hoff(k.hot_off) {}

    __device__ __forceinline__ float& ppath(uint32_t i) { return pp[i * stride]; }

    __device__ __forceinline__ float rnd() {
        uint64_t s1 = t0;
        const uint64_t s0 = t1;
        t0 = s0;
        s1 ^= s1 << 23;
        t1 = s1 ^ s0 ^ (s1 >> 18) ^ (s0 >> 5);
        uint32_t lo = (uint32_t)(t1 + s0);
        return mcx_bits2f(0x3F800000u | (lo >> 9)) - 1.0f;
    }
    __device__ __forceinline__ float scatlen() { return -mcx_logf(rnd() + EPSF); }

    static constexpr bool WORDS = is_word_media<MT>::value;
    static constexpr bool SPLIT = is_split_media<MT>::value;

    __device__ __forceinline__ uint32_t sv_label() const { return svupper ? svup : svlo; }

    /* src/mcx_core.cu:688-729: properties of split-voxel word w (voxel vox)
     * for a photon at q in voxel (i0, i1, i2).  In a split voxel the side
     * of the interface plane decides the label, and the normal is turned to
     * point away from the photon's side. */
    __device__ __forceinline__ void svmc_update(F4& out, uint32_t w, uint32_t vox, const F4& q, int i0, int i1, int i2) {
        if (vox == OUT_MIN || vox == OUT_MAX) {
            out = tab[0];
            return;
        }
        const uint32_t lo = load_media<MT>(kp.media, vox + kp.nvox);
        const uint32_t hi = w & MED_MASK;
        svlo = hi >> 24;
        svup = (hi >> 16) & 0xFFu;
        if (svup) {
            rpx = (float)((hi >> 8) & 0xFFu) * (1.f / 255.f);
            rpy = (float)(hi & 0xFFu) * (1.f / 255.f);
            rpz = (float)(lo >> 24) * (1.f / 255.f);
            rpx += (float)i0;
            rpy += (float)i1;
            rpz += (float)i2;
            nvx = (float)((lo >> 16) & 0xFFu) * (2.f / 255.f) - 1.f;
            nvy = (float)((lo >> 8) & 0xFFu) * (2.f / 255.f) - 1.f;
            nvz = (float)(lo & 0xFFu) * (2.f / 255.f) - 1.f;
            const float r = 1.f / __builtin_sqrtf(nvx * nvx + nvy * nvy + nvz * nvz);
            nvx = nvx * r;
            nvy = nvy * r;
            nvz = nvz * r;
            if ((rpx - q.x) * nvx + (rpy - q.y) * nvy + (rpz - q.z) * nvz < 0.f) {
                out = tab[svup];
                svupper = 1;
                nvx = -nvx;
                nvy = -nvy;
                nvz = -nvz;
            } else {
                out = tab[svlo];
                svupper = 0;
            }
            svsplit = 1;
        } else {
            out = tab[svlo];
            svsplit = 0;
            svupper = 0;
        }
    }

    /* src/mcx_core.cu:747-766: does the step of length len from p cross the
     * interface plane?  If so, shorten the step to the crossing */
    __device__ __forceinline__ bool ray_plane(float& len, float& slen) const {
        if (v.x * nvx + v.y * nvy + v.z * nvz <= 0.f) return false;
        const float p1x = p.x + len * v.x, p1y = p.y + len * v.y, p1z = p.z + len * v.z;
        const float d0 = (p.x - rpx) * nvx + (p.y - rpy) * nvy + (p.z - rpz) * nvz;
        const float d1 = (p1x - rpx) * nvx + (p1y - rpy) * nvy + (p1z - rpz) * nvz;
        if (d0 * d1 > 0.f) return false;
        const float len0 = len * d0 / (d0 - d1);
        len = (len0 > 0.f) ? len0 : len;
        slen = len * prop.y * ((v.w + 1.f > (float)kp.gscatter) ? (1.f - prop.z) : 1.f);
        return true;
    }

    /* src/mcx_core.cu:782-828: Fresnel reflection or refraction at the
     * interface plane; true when the photon transmits into the background */
    __device__ __forceinline__ bool reflectray(float n1) {
        const float Icos = fabsf(v.x * nvx + v.y * nvy + v.z * nvz);
        const float n2 = tab[svupper ? svup : svlo].w;
        const float tmp0 = n1 * n1, tmp1 = n2 * n2;
        float tmp2 = 1.f - tmp0 / tmp1 * (1.f - Icos * Icos);
        if (tmp2 > 0.f) {
            float Re = tmp0 * Icos * Icos + tmp1 * tmp2;
            tmp2 = __builtin_sqrtf(tmp2);
            const float Im = 2.f * n1 * n2 * Icos * tmp2;
            float Rtotal = (Re - Im) / (Re + Im);
            Re = tmp1 * Icos * Icos + tmp0 * tmp2 * tmp2;
            Rtotal = (Rtotal + (Re - Im) / (Re + Im)) * 0.5f;
            if (rnd() <= Rtotal) {
                v.x = v.x + (-2.f * Icos) * nvx;
                v.y = v.y + (-2.f * Icos) * nvy;
                v.z = v.z + (-2.f * Icos) * nvz;
                svupper = !svupper;
            } else {
                v.x = v.x + (-Icos) * nvx;
                v.y = v.y + (-Icos) * nvy;
                v.z = v.z + (-Icos) * nvz;
                const float r = n1 / n2;
                v.x = tmp2 * nvx + r * v.x;
                v.y = tmp2 * nvy + r * v.y;
                v.z = tmp2 * nvz + r * v.z;
                nvx = -nvx;
                nvy = -nvy;
                nvz = -nvz;
                /* the reference tests the side flag itself here, not the
                 * upper label: only a lower region of label 0 ends the photon */
                if ((svupper ? svupper : svlo) == 0) return true;
                prop = tab[svupper ? svup : svlo];
            }
        } else {
            v.x = v.x + (-2.f * Icos) * nvx;
            v.y = v.y + (-2.f * Icos) * nvy;
            v.z = v.z + (-2.f * Icos) * nvz;
            svupper = !svupper;
        }
        const float r = 1.f / __builtin_sqrtf(v.x * v.x + v.y * v.y + v.z * v.z);
        v.x = v.x * r;
        v.y = v.y * r;
        v.z = v.z * r;
        calc_rv();
        return false;
    }

    /* src/mcx_core.cu:582-687: optical properties of voxel word m (voxel
     * vox) for the non-label formats.  Fields a format does not encode keep
     * their value in q (a lane's properties start as row 1 of the table).
     * The mixed-label format draws one random number and narrows m to the
     * label drawn, so later decodes of the same word keep that label. */
    __device__ __forceinline__ void decode(F4& q, uint32_t& m, uint32_t vox) {
        const uint32_t w = m & MED_MASK;
        const uint32_t fmt = kp.mediaformat;
        if (fmt == MCX_MEDIA_LABEL_HALF) {
            q = tab[w & 0x3FFFu];
            const float h = fabsf(half_to_float(w >> 16));
            const uint32_t k = (w >> 14) & 3u;
            if (k == 0) q.x = h;
            else if (k == 1) q.y = h;
            else if (k == 2) q.z = h;
            else q.w = h;
        } else if (fmt == MCX_MEDIA_MUA_FLOAT) {
            q.x = fabsf(mcx_bits2f(m));
            q.w = tab[w != 0u].w;
        } else if (fmt == MCX_MEDIA_AS_F2H || fmt == MCX_MEDIA_AS_HALF) {
            q.x = fabsf(half_to_float(w & 0xFFFFu));
            q.y = fabsf(half_to_float(w >> 16));
            q.w = tab[w != 0u].w;
        } else if (fmt == MCX_MEDIA_ASGN_F2H) {
            if (vox == OUT_MIN || vox == OUT_MAX) {
                q = tab[0];
                return;
            }
            const uint32_t w2 = load_media<MediaWord>(kp.media, vox + kp.nvox);
            q.x = fabsf(half_to_float(w & 0xFFFFu));
            q.y = fabsf(half_to_float(w >> 16));
            q.z = fabsf(half_to_float(w2 & 0xFFFFu));
            q.w = fabsf(half_to_float(w2 >> 16));
        } else if (fmt == MCX_MEDIA_2LABEL_MIX) {
            const uint32_t frac = w >> 16;
            if (frac > 0u) {
                if (rnd() * 32767.f < (float)frac) {
                    q = tab[(w >> 8) & 0xFFu];
                    m >>= 8;
                } else {
                    q = tab[w & 0xFFu];
                }
                m &= 0xFFFFu;
            } else {
                q = tab[w & 0xFFu];
            }
        } else if (fmt == MCX_MEDIA_ASGN_BYTE) {
            const F4 a = tab[1], b = tab[2];
            q.x = (float)(w & 0xFFu) * (1.f / 255.f) * (b.x - a.x) + a.x;
            q.y = (float)((w >> 8) & 0xFFu) * (1.f / 255.f) * (b.y - a.y) + a.y;
            q.z = (float)((w >> 16) & 0xFFu) * (1.f / 255.f) * (b.z - a.z) + a.z;
            q.w = (float)(w >> 24) * (1.f / 127.f) * (b.w - a.w) + a.w;
        } else if (fmt == MCX_MEDIA_AS_SHORT) {
            const F4 a = tab[1], b = tab[2];
            q.x = (float)(w & 0xFFFFu) * (1.f / 65535.f) * (b.x - a.x) + a.x;
            q.y = (float)(w >> 16) * (1.f / 65535.f) * (b.y - a.y) + a.y;
            q.w = tab[w != 0u].w;
        }
    }

    __device__ __forceinline__ void setprop(uint32_t& m, uint32_t vox) {
        if (SPLIT)
            svmc_update(prop, m, vox, p, fd0, fd1, fd2);
        else if (WORDS)
            decode(prop, m, vox);
        else
            prop = tab[m & MED_MASK];
    }
    __device__ __forceinline__ void calc_rv() {
        rv0 = 1.f / v.x;
        rv1 = 1.f / v.y;
        rv2 = 1.f / v.z;
    }

    /* src/mcx_core.cu:493-507 */
    __device__ __forceinline__ float hitgrid(const F4& q, int& i0, int& i1, int& i2, int& face, float r0, float r1, float r2) const {
        float h0 = fabsf(((float)(i0 + (v.x > 0.f)) - q.x) * r0);
        float h1 = fabsf(((float)(i1 + (v.y > 0.f)) - q.y) * r1);
        float h2 = fabsf(((float)(i2 + (v.z > 0.f)) - q.z) * r2);
        float dist = fminf(fminf(h0, h1), h2);
        face = (dist == h0 ? 0 : (dist == h1 ? 1 : 2));
        return dist;
    }

    /* src/mcx_core.cu:1011-1031 */
    __device__ __forceinline__ void rotatevector(float stheta, float ctheta, float sphi, float cphi) {
        if (v.z > -1.f + EPSF && v.z < 1.f - EPSF) {
            float tmp0 = 1.f - v.z * v.z;
            float tmp1 = stheta / __builtin_sqrtf(tmp0);
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
        float tmp0 = 1.f / __builtin_sqrtf(v.x * v.x + v.y * v.y + v.z * v.z);
        v.x *= tmp0; v.y *= tmp0; v.z *= tmp0;
    }

    /* src/mcx_core.cu:970-999 */
    __device__ __forceinline__ void rotatevector2d(float stheta, float ctheta) {
        if (kp.is2d == 1) {
            float ny = v.y * ctheta - v.z * stheta, nz = v.y * stheta + v.z * ctheta;
            v.x = 0.f; v.y = ny; v.z = nz;
        } else if (kp.is2d == 2) {
            float nx = v.x * ctheta - v.z * stheta, nz = v.x * stheta + v.z * ctheta;
            v.x = nx; v.y = 0.f; v.z = nz;
        } else {
            float nx = v.x * ctheta - v.y * stheta, ny = v.x * stheta + v.y * ctheta;
            v.x = nx; v.y = ny; v.z = 0.f;
        }
        float tmp0 = 1.f / __builtin_sqrtf(v.x * v.x + v.y * v.y + v.z * v.z);
        v.x *= tmp0; v.y *= tmp0; v.z *= tmp0;
    }

    /* component `i` of v, selected with bit masks: a select between
     * struct members would become a dynamically indexed load and keep the
     * whole lane state in scratch memory */
    __device__ __forceinline__ float vcomp(int i) const {
        const uint32_t mx = 0u - (uint32_t)(i == 0), my = 0u - (uint32_t)(i == 1), mz = ~(mx | my);
        return mcx_bits2f((mcx_f2bits(v.x) & mx) | (mcx_f2bits(v.y) & my) | (mcx_f2bits(v.z) & mz));
    }

    /* src/mcx_core.cu:523-537 */
    __device__ __forceinline__ void transmit(float n1, float n2, int face) {
        float tmp0 = n1 / n2;
        v.x *= tmp0; v.y *= tmp0; v.z *= tmp0;
        if (face == 0) {
            tmp0 = v.y * v.y + v.z * v.z;
            v.x = (tmp0 < 1.f) ? __builtin_sqrtf(1.f - tmp0) * sgnf(v.x) : 0.f;
        } else if (face == 1) {
            tmp0 = v.x * v.x + v.z * v.z;
            v.y = (tmp0 < 1.f) ? __builtin_sqrtf(1.f - tmp0) * sgnf(v.y) : 0.f;
        } else {
            tmp0 = v.x * v.x + v.y * v.y;
            v.z = (tmp0 < 1.f) ? __builtin_sqrtf(1.f - tmp0) * sgnf(v.z) : 0.f;
        }
        tmp0 = 1.f / __builtin_sqrtf(v.x * v.x + v.y * v.y + v.z * v.z);
        v.x *= tmp0; v.y *= tmp0; v.z *= tmp0;
    }

    /* src/mcx_core.cu:552-570 */
    __device__ __forceinline__ float reflectcoeff(float n1, float n2, int face) const {
        float Icos = fabsf(vcomp(face));
        float tmp0 = n1 * n1;
        float tmp1 = n2 * n2;
        float tmp2 = 1.f - tmp0 / tmp1 * (1.f - Icos * Icos);
        if (tmp2 > 0.f) {
            float Re = tmp0 * Icos * Icos + tmp1 * tmp2;
            tmp2 = __builtin_sqrtf(tmp2);
            float Im = 2.f * n1 * n2 * Icos * tmp2;
            float Rtotal = (Re - Im) / (Re + Im);
            Re = tmp1 * Icos * Icos + tmp0 * tmp2 * tmp2;
            Rtotal = (Rtotal + (Re - Im) / (Re + Im)) * 0.5f;
            return Rtotal;
        }
        return 1.f;
    }

    /* src/mcx_core.cu:258-270 */
    __device__ __forceinline__ uint32_t finddetector() const {
        for (uint32_t i = kp.maxmedia + 1; i < kp.maxmedia + kp.detnum + 1; i++) {
            const F4 d = tab[i];
            if ((d.x - p.x) * (d.x - p.x) + (d.y - p.y) * (d.y - p.y) + (d.z - p.z) * (d.z - p.z) < d.w * d.w)
                return (d.w < 0.f) ? 0 : i - kp.maxmedia;
        }
        return 0;
    }

    /* src/mcx_core.cu:365-417, slot reservation aggregated per wave */
    __device__ __forceinline__ void savedet(bool want, uint32_t isdet) {
        int detid = 0;
        if (want) detid = (isdet == OUT_MIN) ? -1 : (int)finddetector();
        const bool hit = want && detid != 0;
        const uint64_t mask = __ballot(hit);
        if (mask == 0) return;
        const uint32_t lane = __lane_id();
        const uint32_t leader = __ffsll((unsigned long long)mask) - 1;
        uint32_t base = 0;
        if (lane == leader) base = atom_add_u32(kp.detcount, (uint32_t)__popcll(mask));
        base = __shfl(base, leader);
        if (!hit) return;
        const uint32_t slot = base + (uint32_t)__popcll(mask & ((1ull << lane) - 1ull));
        if (slot < kp.maxdetphoton) {
            MCX_GLB float* rec = glb(kp.detdata) + (size_t)slot * kp.reclen;
            uint32_t k = 0;
            const uint32_t flag = kp.savedetflag;
            if (flag & 1) rec[k++] = (float)detid;
            for (uint32_t i = 0; i < kp.partialdata; i++) rec[k++] = ppath(i);
            if ((flag >> 4) & 1) { rec[k++] = p.x; rec[k++] = p.y; rec[k++] = p.z; }
            if ((flag >> 5) & 1) { rec[k++] = v.x; rec[k++] = v.y; rec[k++] = v.z; }
            if ((flag >> 6) & 1) rec[k++] = winit;
            if (kp.saveseed) {
                MCX_GLB uint32_t* sd = glb(kp.seeddata) + (size_t)slot * 4;
                for (uint32_t i = 0; i < 4; i++) sd[i] = mcx_f2bits(ppath(kp.partialdata + i));
            }
        } else if (kp.savedet == 3) {
            atom_add_u32(kp.detcount, ~0u);
        }
    }

    /* src/mcx_core.cu:866-956 */
    __device__ __forceinline__ int skipvoid() {
        int count = 1, i0, i1, i2, face = -1;
        i0 = (int)floorf(p.x); i1 = (int)floorf(p.y); i2 = (int)floorf(p.z);
        while (1) {
            if ((uint32_t)i0 < kp.nx && (uint32_t)i1 < kp.ny && (uint32_t)i2 < kp.nz) {
                uint32_t id = i2 * kp.dimxy + i1 * kp.nx + i0;
                if (load_media<MT>(kp.media, id) & MED_MASK) {
                    p.x -= v.x; p.y -= v.y; p.z -= v.z;
                    i0 = (int)floorf(p.x); i1 = (int)floorf(p.y); i2 = (int)floorf(p.z);
                    f.y -= kp.minaccumtime;
                    id = i2 * kp.dimxy + i1 * kp.nx + i0;
                    count = 0;
                    while (!((uint32_t)i0 < kp.nx && (uint32_t)i1 < kp.ny && (uint32_t)i2 < kp.nz) ||
                           !(load_media<MT>(kp.media, id) & MED_MASK)) {
                        float dist = hitgrid(p, i0, i1, i2, face, rv0, rv1, rv2);
                        f.y += kp.minaccumtime * dist;
                        p.x = p.x + dist * v.x; p.y = p.y + dist * v.y; p.z = p.z + dist * v.z;
                        if (face == 0) i0 += (v.x > 0.f ? 1 : -1);
                        if (face == 1) i1 += (v.y > 0.f ? 1 : -1);
                        if (face == 2) i2 += (v.z > 0.f ? 1 : -1);
                        id = i2 * kp.dimxy + i1 * kp.nx + i0;
                        if (count++ > 3) break;
                    }
                    /* the entry search gave up outside the grid: no tissue
                     * entry (the reference would read an out-of-range label) */
                    if (!((uint32_t)i0 < kp.nx && (uint32_t)i1 < kp.ny && (uint32_t)i2 < kp.nz)) return -1;
                    f.y = kp.voidtime ? f.y : 0.f;
                    F4 h;
                    if (SPLIT) {
                        svmc_update(h, load_media<MT>(kp.media, id), id, p, i0, i1, i2);
                    } else if (WORDS) {
                        /* src/mcx_core.cu:921-922 (the word is decoded into a
                         * local: the volume stays read-only) */
                        uint32_t w = load_media<MT>(kp.media, id);
                        h = tab[1];
                        decode(h, w, id);
                    } else {
                        h = tab[load_media<MT>(kp.media, id) & MED_MASK];
                    }
                    if (kp.isspecular && h.w != tab[0].w) {
                        p.w *= 1.f - reflectcoeff(tab[0].w, h.w, face);
                        if (p.w > EPSF) transmit(tab[0].w, h.w, face);
                    }
                    return (int)id;
                }
            }
            if ((p.x < 0.f && v.x <= 0.f) || (p.x >= kp.maxidx_x && v.x >= 0.f) ||
                (p.y < 0.f && v.y <= 0.f) || (p.y >= kp.maxidx_y && v.y >= 0.f) ||
                (p.z < 0.f && v.z <= 0.f) || (p.z >= kp.maxidx_z && v.z >= 0.f))
                return -1;
            p.x = p.x + v.x; p.y = p.y + v.y; p.z = p.z + v.z;
            i0 = (int)floorf(p.x); i1 = (int)floorf(p.y); i2 = (int)floorf(p.z);
            f.y += kp.minaccumtime;
            if (count++ > (int)kp.maxvoidstep) return -1;
        }
    }

    __device__ __forceinline__ void locate_launch() {
        idx1d = (uint32_t)((int)floorf(p.z) * (int)kp.dimxy + (int)floorf(p.y) * (int)kp.nx + (int)floorf(p.x));
        if (p.x < 0.f || p.y < 0.f || p.z < 0.f || p.x >= kp.maxidx_x || p.y >= kp.maxidx_y || p.z >= kp.maxidx_z)
            mediaid = 0;
        else
            mediaid = load_media<MT>(kp.media, idx1d);
    }

    /* source sampling for non-pencil sources (src/mcx_core.cu:1261-1641) */
    __device__ __forceinline__ void launch_source(bool& skip) {
        const float* s1 = kp.srcparam1;
        const float* s2 = kp.srcparam2;
        int canfocus = 1;
        float ox = kp.srcpos[0], oy = kp.srcpos[1], oz = kp.srcpos[2];
        skip = false;
        switch (kp.srctype) {
            case 4: case 5: case 6: case 14: case 15: {
                float rx = rnd();
                float ry = rnd();
                if (kp.srctype == 15) {
                    const float rz = rnd();
                    p.x = p.x + rx * s1[0];
                    p.y = p.y + ry * s1[1];
                    p.z = p.z + rz * s1[2];
                    /* src/mcx_core.cu:1298-1302, one pattern (srcnum 1) */
                    p.w = kp.srcpos[3] * glb(kp.srcpattern)[(int)(rz * JUST_BELOW_ONE * s1[2]) * (int)s1[1] * (int)s1[0] +
                                                              (int)(ry * JUST_BELOW_ONE * s1[1]) * (int)s1[0] +
                                                              (int)(rx * JUST_BELOW_ONE * s1[0])];
                } else {
                    p.x = p.x + rx * s1[0] + ry * s2[0];
                    p.y = p.y + rx * s1[1] + ry * s2[1];
                    p.z = p.z + rx * s1[2] + ry * s2[2];
                }
                if (kp.srctype == 5) {
                    /* src/mcx_core.cu:1284-1287, one pattern (srcnum 1) */
                    p.w = kp.srcpos[3] *
                          glb(kp.srcpattern)[(int)(ry * JUST_BELOW_ONE * s2[3]) * (int)s1[3] + (int)(rx * JUST_BELOW_ONE * s1[3])];
                } else if (kp.srctype == 6) {
                    float s, c;
                    float ang = (floorf(s1[3]) * rx + floorf(s2[3]) * ry + s1[3] - floorf(s1[3]));
                    mcx_sincos2pi(ang, &s, &c);
                    p.w = kp.srcpos[3] * (c * (1.f - s2[3] + floorf(s2[3])) + 1.f) * 0.5f;
                } else if (kp.srctype == 14) {
                    p.x = kp.srcpos[0] + floorf(rx * s1[3]) * s1[0] / (s1[3] - 1.f) + floorf(ry * s2[3]) * s2[0] / (s2[3] - 1.f);
                    p.y = kp.srcpos[1] + floorf(rx * s1[3]) * s1[1] / (s1[3] - 1.f) + floorf(ry * s2[3]) * s2[1] / (s2[3] - 1.f);
                    p.z = kp.srcpos[2] + floorf(rx * s1[3]) * s1[2] / (s1[3] - 1.f) + floorf(ry * s2[3]) * s2[2] / (s2[3] - 1.f);
                }
                locate_launch();
                ox = ox + (s1[0] + s2[0]) * 0.5f;
                oy = oy + (s1[1] + s2[1]) * 0.5f;
                oz = oz + (s1[2] + s2[2]) * 0.5f;
                break;
            }
            case 9: case 10: {
                /* src/mcx_core.cu:1335-1366 */
                const float rx = rnd();
                const float ry = rnd();
                float v2w = s1[3] * (1.f / __builtin_sqrtf(s1[0] * s1[0] + s1[1] * s1[1] + s1[2] * s1[2]));
                const float v2x = v2w * (kp.srcdir[1] * s1[2] - kp.srcdir[2] * s1[1]);
                const float v2y = v2w * (kp.srcdir[2] * s1[0] - kp.srcdir[0] * s1[2]);
                const float v2z = v2w * (kp.srcdir[0] * s1[1] - kp.srcdir[1] * s1[0]);
                p.x = p.x + rx * s1[0] + ry * v2x;
                p.y = p.y + rx * s1[1] + ry * v2y;
                p.z = p.z + rx * s1[2] + ry * v2z;
                if (kp.srctype == 10) {
                    float sa, ca, sb, cb;
                    mcx_sincos2pi(s2[0] * rx + s2[2], &sa, &ca);
                    mcx_sincos2pi(s2[1] * ry + s2[3], &sb, &cb);
                    p.w = kp.srcpos[3] * (sa * sb + 1.f) * 0.5f;
                } else {
                    float sa, ca;
                    mcx_sincos2pi(s2[0] * rx + s2[1] * ry + s2[2], &sa, &ca);
                    p.w = kp.srcpos[3] * (ca * (1.f - s2[3]) + 1.f) * 0.5f;
                }
                locate_launch();
                ox = ox + (s1[0] + v2x) * 0.5f;
                oy = oy + (s1[1] + v2y) * 0.5f;
                oz = oz + (s1[2] + v2z) * 0.5f;
                break;
            }
            case 11: {
                /* src/mcx_core.cu:1446-1455 */
                float sphi, cphi, stheta, ctheta;
                mcx_sincos2pi(rnd(), &sphi, &cphi);
                const float ang = __builtin_sqrtf(2.f * scatlen()) * (1.f - 2.f * rnd()) * s1[0];
                mcx_sincos2pi(ang * 0.159154943091895336f, &stheta, &ctheta);
                rotatevector(stheta, ctheta, sphi, cphi);
                canfocus = 0;
                break;
            }
            case 16: {
                /* src/mcx_core.cu:1457-1505 */
                float sphi, cphi;
                mcx_sincos2pi(rnd(), &sphi, &cphi);
                float r = __builtin_sqrtf(0.5f * scatlen()) * s1[0];
                const float ax = -s1[1] / s1[2];
                const float ay = 1.f / __builtin_sqrtf(r * r + s1[2] * s1[2]);
                float px = r * (cphi - ax * sphi), py = r * (sphi + ax * cphi), pz = 0.f;
                const float dx = -r * sphi * ay, dy = r * cphi * ay, dz = s1[2] * ay;
                if (v.z > -1.f + EPSF && v.z < 1.f - EPSF) {
                    r = 1.f - v.z * v.z;
                    const float st = __builtin_sqrtf(r);
                    r = 1.f / __builtin_sqrtf(r);
                    const float cp = v.x * r, sp = v.y * r;
                    const float qx = px * cp * v.z - py * sp, qy = px * sp * v.z + py * cp, qz = -px * st;
                    px = qx; py = qy; pz = qz;
                    const float nx = dx * cp * v.z - dy * sp + dz * cp * st;
                    const float ny = dx * sp * v.z + dy * cp + dz * sp * st;
                    const float nz = -dx * st + dz * v.z;
                    v.x = nx; v.y = ny; v.z = nz;
                } else {
                    v.z = (v.z > 0.f) ? dz : -dz;
                    v.x = dx; v.y = dy;
                }
                p.x = px + kp.srcpos[0];
                p.y = py + kp.srcpos[1];
                p.z = pz + kp.srcpos[2];
                locate_launch();
                canfocus = 0;
                break;
            }
            case 12: case 13: {
                /* src/mcx_core.cu:1508-1568 */
                float r = rnd();
                p.x = p.x + r * s1[0];
                p.y = p.y + r * s1[1];
                p.z = p.z + r * s1[2];
                if (kp.srctype == 12) {
                    float sphi, cphi;
                    r = 1.f / __builtin_sqrtf(s1[0] * s1[0] + s1[1] * s1[1] + s1[2] * s1[2]);
                    if (s2[0] > 0.f) {
                        const float ax = s1[0] * r, ay = s1[1] * r, az = s1[2] * r;
                        r = v.x * ax + v.y * ay + v.z * az;
                        v.x = v.x - r * ax; v.y = v.y - r * ay; v.z = v.z - r * az;
                        r = 1.f / __builtin_sqrtf(v.x * v.x + v.y * v.y + v.z * v.z);
                        v.x = v.x * r; v.y = v.y * r; v.z = v.z * r;
                        r = s2[0] * (2.f * rnd() - 1.f);
                        mcx_sincos2pi(r * 0.159154943091895336f, &sphi, &cphi);
                        const float cx = ay * v.z - az * v.y, cy = az * v.x - ax * v.z, cz = ax * v.y - ay * v.x;
                        v.x = v.x * cphi + cx * sphi;
                        v.y = v.y * cphi + cy * sphi;
                        v.z = v.z * cphi + cz * sphi;
                    } else {
                        v.x = s1[0] * r; v.y = s1[1] * r; v.z = s1[2] * r;
                        mcx_sincos2pi(rnd(), &sphi, &cphi);
                        rotatevector(1.f, 0.f, sphi, cphi);
                    }
                } else if (s2[0] > 0.f || s2[1] > 0.f) {
                    float sphi, cphi;
                    mcx_sincos2pi(rnd(), &sphi, &cphi);
                    r = __builtin_sqrtf(2.f * scatlen());
                    cphi *= s2[0] * r;
                    sphi *= s2[1] * r;
                    sphi *= 1.f / __builtin_sqrtf(s1[0] * s1[0] + s1[1] * s1[1] + s1[2] * s1[2]);
                    const float cx = s1[1] * v.z - s1[2] * v.y, cy = s1[2] * v.x - s1[0] * v.z, cz = s1[0] * v.y - s1[1] * v.x;
                    cphi *= 1.f / __builtin_sqrtf(cx * cx + cy * cy + cz * cz);
                    v.x += cphi * cx + sphi * s1[0];
                    v.y += cphi * cy + sphi * s1[1];
                    v.z += cphi * cz + sphi * s1[2];
                    r = 1.f / __builtin_sqrtf(v.x * v.x + v.y * v.y + v.z * v.z);
                    v.x *= r; v.y *= r; v.z *= r;
                }
                locate_launch();
                ox = kp.srcpos[0] + s1[0] * 0.5f;
                oy = kp.srcpos[1] + s1[1] * 0.5f;
                oz = kp.srcpos[2] + s1[2] * 0.5f;
                canfocus = (kp.srctype == 13);
                break;
            }
            case 8: case 17: case 3: {
                float sphi, cphi, r;
                if (kp.srctype == 17 && (s1[2] > 0.f || s1[3] > 0.f)) {
                    float phi = fabsf(s1[2] - s1[3]) * rnd() + fminf(s1[2], s1[3]);
                    mcx_sincos2pi(phi * 0.159154943091895336f, &sphi, &cphi);
                } else {
                    mcx_sincos2pi(rnd(), &sphi, &cphi);
                }
                if (kp.srctype == 8 || kp.srctype == 17) {
                    r = __builtin_sqrtf(rnd() * fabsf(s1[0] * s1[0] - s1[1] * s1[1]) + s1[1] * s1[1]);
                } else if (fabsf(kp.srcdir[3]) < 1e-5f || fabsf(s1[1]) < 1e-5f) {
                    r = __builtin_sqrtf(0.5f * scatlen()) * s1[0];
                } else {
                    r = s1[0] * s1[0] * 3.14159265358979f / s1[1];
                    r = __builtin_sqrtf(0.5f * scatlen() * (1.f + (kp.srcdir[3] * kp.srcdir[3] / (r * r)))) * s1[0];
                }
                if (v.z > -1.f + EPSF && v.z < 1.f - EPSF) {
                    float tmp0 = 1.f - v.z * v.z;
                    float tmp1 = r / __builtin_sqrtf(tmp0);
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
            case 1: case 2: case 7: {
                float sphi, cphi, stheta, ctheta;
                mcx_sincos2pi(rnd(), &sphi, &cphi);
                if (kp.srctype == 2) {
                    if (s1[1] > 0.f) {
                        float ang = rnd() * s1[0];
                        mcx_sincos2pi(ang * 0.159154943091895336f, &stheta, &ctheta);
                    } else {
                        float cmax, dummy;
                        mcx_sincos2pi(s1[0] * 0.159154943091895336f, &dummy, &cmax);
                        ctheta = rnd() * (1.f - cmax) + cmax;
                        stheta = mcx_sin_from_cos(ctheta);
                    }
                } else if (kp.srctype == 1) {
                    ctheta = 2.f * rnd() - 1.f;
                    stheta = mcx_sin_from_cos(ctheta);
                } else {
                    mcx_sincos2pi(rnd() * 0.5f, &stheta, &ctheta);
                }
                rotatevector(stheta, ctheta, sphi, cphi);
                canfocus = 0;
                break;
            }
            default:
                break;
        }
        if (fabsf(p.w) <= kp.minenergy) {
            skip = true;
            return;
        }
        if (canfocus) {
            const float fw = kp.srcdir[3];
            if (__builtin_isnan(fw)) {
                float sphi, cphi;
                mcx_sincos2pi(rnd(), &sphi, &cphi);
                float ctheta = 2.f * rnd() - 1.f;
                rotatevector(mcx_sin_from_cos(ctheta), ctheta, sphi, cphi);
            } else if (fw < 0.f && __builtin_isinf(fw)) {
                float sphi, cphi;
                mcx_sincos2pi(rnd(), &sphi, &cphi);
                float stheta = __builtin_sqrtf(rnd());
                float ctheta = __builtin_sqrtf(1.f - stheta * stheta);
                rotatevector(stheta, ctheta, sphi, cphi);
            } else if (fw != 0.f) {
                float Rn2 = (float)((fw > 0.f) - (fw < 0.f));
                ox += fw * v.x;
                oy += fw * v.y;
                oz += fw * v.z;
                v.x = Rn2 * (ox - p.x);
                v.y = Rn2 * (oy - p.y);
                v.z = Rn2 * (oz - p.z);
                Rn2 = 1.f / __builtin_sqrtf(v.x * v.x + v.y * v.y + v.z * v.z);
                v.x *= Rn2; v.y *= Rn2; v.z *= Rn2;
            }
        }
    }

    /* src/mcx_core.cu:1089-1717: terminate the current photon, launch the
     * next one.  Returns 0 launched, 1 lane finished, -1 launch failed. */
    __device__ __forceinline__ int launch(uint32_t isdet) {
        w0 = 1.f;
        if (kp.savedet == 3 && __hip_atomic_load(glb(kp.detcount), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) >= kp.maxdetphoton)
            return 1;
        const bool live = fabsf(p.w) >= 0.f;
        if (live) esc += p.w;
        /* src/mcx_core.cu:1124-1145: diffuse reflectance, the escaping weight
         * tallied with a negative sign in the void voxel entered */
        if (kp.issaveref && mediaid == 0 && idx1d != OUT_MIN && idx1d != OUT_MAX && p.w > 0.f) {
            const int ts = max(0, min((int)kp.maxgate - 1, (int)floorf((f.y - kp.twin0) * kp.Rtstep)));
            deposit(idx1d, ts, -p.w);
        }
        if (kp.savedet) {
            savedet(live && (isdet & DET_MASK) == DET_MASK && (mediaid == 0 || (SPLIT && sv_label() == 0)), isdet);
            for (uint32_t i = 0; i < kp.partialdata; i++) ppath(i) = 0.f;
        }
        if ((int)f.w >= nphoton - 1) {
            /* With kp.refill a lane runs virtual thread after virtual
             * thread: once one has launched all its photons the lane takes
             * the next unstarted one from a global counter (detcount[3]),
             * so lanes keep stepping until every virtual thread has started
             * instead of idling while the rest of their wave finishes.
             * Each virtual thread keeps its own RNG stream, photon count
             * and energy sums, so the results are those of a launch with
             * one lane per virtual thread. */
            if (!kp.refill) return 1;
            glb(kp.genergy)[2 * idx] = esc;
            glb(kp.genergy)[2 * idx + 1] = launched;
            const uint32_t next = kp.nlanes + atom_add_u32(kp.detcount + 3, 1u);
            if (next >= kp.nvthreads) return 1;
            if (TALLY == MCX_TALLY_F64 && kp.nmaps_mask) {
                /* the open block belongs to the old run's tally */
                block_finish();
                tally64 = kp.field64 + (size_t)(next & kp.nmaps_mask) * kp.mapstride;
            }
            idx = next;
            nphoton = kp.threadphoton + ((int)idx < kp.oddphotons);
            esc = glb(kp.genergy)[2 * idx];
            launched = glb(kp.genergy)[2 * idx + 1];
            const MCX_GLB uint32_t* sd = glb(kp.seeds) + 4 * idx;
            t0 = ((uint64_t)sd[0] << 32) | sd[1];
            t1 = ((uint64_t)sd[2] << 32) | sd[3];
            f.w = -1.f;
            for (uint32_t i = 0; i < kp.partialdata; i++) ppath(i) = 0.f;
        }
        /* src/mcx_core.cu:1205-1215: a replayed photon restarts from its
         * saved RNG state; with seed saving the launch state is kept (in
         * this lane's LDS column after the partial paths) until detection */
        gid = idx * (uint32_t)kp.threadphoton +
              (kp.replayindex ? (uint32_t)min((int)idx, kp.oddphotons - 1) : min(idx, (uint32_t)kp.oddphotons)) +
              (uint32_t)((int)f.w + 1);
        if (kp.replay) {
            const MCX_GLB uint32_t* rs = glb(kp.replayseed) + (size_t)gid * 4;
            t0 = ((uint64_t)rs[1] << 32) | rs[0];
            t1 = ((uint64_t)rs[3] << 32) | rs[2];
        }
        if (kp.saveseed) {
            ppath(kp.partialdata + 0) = mcx_bits2f((uint32_t)t0);
            ppath(kp.partialdata + 1) = mcx_bits2f((uint32_t)(t0 >> 32));
            ppath(kp.partialdata + 2) = mcx_bits2f((uint32_t)t1);
            ppath(kp.partialdata + 3) = mcx_bits2f((uint32_t)(t1 >> 32));
        }
        do {
            p = {kp.srcpos[0], kp.srcpos[1], kp.srcpos[2], kp.srcpos[3]};
            v = {kp.srcdir[0], kp.srcdir[1], kp.srcdir[2], kp.srcdir[3]};
            f = {0.f, 0.f, kp.minaccumtime, f.w};
            idx1d = mcx_f2bits(kp.srcparam2[2]);
            mediaid = mcx_f2bits(kp.srcparam2[3]);
            if (SPLIT) svsplit = svlo = svup = svupper = 0;
            /* ISPENCIL 2: decided per launch parameter block (batched launch) */
            if (!(ISPENCIL == 1 || (ISPENCIL == 2 && kp.srctype == 0))) {
                bool skip;
                launch_source(skip);
                if (skip) continue;
            }
            calc_rv();
            if ((mediaid & MED_MASK) == 0) {
                int i1 = skipvoid();
                if (i1 >= 0) {
                    idx1d = (uint32_t)i1;
                    mediaid = load_media<MT>(kp.media, idx1d);
                }
            }
            fd0 = (int)floorf(p.x); fd1 = (int)floorf(p.y); fd2 = (int)floorf(p.z);
            w0 += 1.f;
            if (w0 > (float)kp.maxvoidstep) return -1;
        } while ((mediaid & MED_MASK) == 0 || fabsf(p.w) <= kp.minenergy);
        f.w += 1.f;
        setprop(mediaid, idx1d);
        launched += p.w;
        w0 = p.w;
        winit = p.w;
        v.w = EPSF;
        f.z = 0.f;
        nlaunch++;
        pstep = 0;
        return 0;
    }

    /* Second half of an FP32 deposit.  Shadow mode (src/mcx_core.cu:
     * 2246-2253): once the returned old value exceeds MAX_ACCUM, move
     * MAX_ACCUM from the value half to the shadow half.  Hi/lo mode: the
     * atomic stored s = fl(old + w); TwoSum gives the exact rounding error,
     * which goes to the lo half, and whole chunks leave the hi half
     * (hilo_complete).  The moves preserve the merged sum, so doing them
     * one deposit later changes only when, not what. */
    __device__ __forceinline__ void flush_pending() {
        if (BLOCKED) block_finish();
        if (TALLY == MCX_TALLY_F64 || TALLY == MCX_TALLY_F64_XCD || pend_i == ~(size_t)0) return;
        const size_t i = pend_i;
        const float a = pend_old, w = pend_w;
        pend_i = ~(size_t)0;
        if (TALLY == MCX_TALLY_F32_SHADOW) {
            if (fabsf(a) > MAX_ACCUM) {
                atom_add_nr(kp.field + i, (a > 0.f) ? -MAX_ACCUM : MAX_ACCUM);
                atom_add_nr(kp.field2 + i, (a > 0.f) ? MAX_ACCUM : -MAX_ACCUM);
            }
        } else {
            hilo_complete(kp.field + i, kp.field2 + i, kp.fieldk + i, a, w);
        }
    }

    /* the wave adds the blocks its lanes are leaving (pend_i: the cell of
     * the deposit that did not fit a lane's block) to the tally, 8 lanes to
     * a block; then those lanes open the block of their new deposit.
     * Called where every lane still stepping is active. */
    __device__ __forceinline__ void block_flush() {
        if (!BLOCKED || !kp.tblock) return;
        const bool mine = pend_i != ~(size_t)0;
        const uint64_t m = __ballot(mine);
        if (m == 0) return;
        extern __shared__ F4 lds[];
        const float* all = (const float*)((const char*)lds + kp.blk_off);
        uint32_t* q = (uint32_t*)((char*)lds + kp.bq_off) + (threadIdx.x & ~63u);
        const uint64_t act = __ballot(true);
        const uint32_t na = (uint32_t)__popcll(act), rank = lanes_below(act);
        const uint32_t work = 8u * (uint32_t)__popcll(m);
        /* whole blocks per instruction: a multiple of 8 workers */
        const uint32_t nw = na >= 8u ? (na & ~7u) : na;
        if (mine) q[lanes_below(m)] = threadIdx.x;
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        __builtin_amdgcn_wave_barrier();
        if (rank < nw) {
            for (uint32_t item = rank; item < work; item += nw) {
                const float* b = all + 9u * q[item >> 3];
                const float w = b[item & 7u];
                if (w != 0.f) atom_add_nr(tally64 + (((size_t)mcx_f2bits(b[8]) << 3) + (item & 7u)), (double)w);
            }
        }
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        __builtin_amdgcn_wave_barrier();
        if (mine) {
            nreq++;
            for (uint32_t c = 0; c < 8; c++) blk[c] = 0.f;
            blk[pend_i & 7u] = pend_w;
            bopen = (uint32_t)(pend_i >> 3);
            blk[8] = mcx_bits2f(bopen);
            pend_i = ~(size_t)0;
        }
    }

    /* a lane leaving the kernel adds its block (and a deposit still
     * waiting for one) to the tally by itself */
    __device__ __forceinline__ void block_finish() {
        if (!BLOCKED || !kp.tblock) return;
        if (bopen != MCX_NO_BLOCK)
            for (uint32_t c = 0; c < 8; c++)
                if (blk[c] != 0.f) {
                    atom_add_nr(tally64 + (((size_t)bopen << 3) + c), (double)blk[c]);
                    nreq++;
                }
        if (pend_i != ~(size_t)0) {
            atom_add_nr(tally64 + pend_i, (double)pend_w);
            nreq++;
        }
        bopen = MCX_NO_BLOCK;
        pend_i = ~(size_t)0;
    }

    /* `main`: the voxel-exit deposit of the step loop (at most one per
     * loop iteration), which may wait in the lane's block.  (cx, cy, cz):
     * the caller's coordinates of `vox` if it has them (the step loop's
     * voxel left); they are used only when they index `vox` inside the
     * grid, otherwise the index is divided back into coordinates. */
    __device__ __forceinline__ void deposit(uint32_t vox, int tshift, float weight, bool main = false,
                                            uint32_t cx = ~0u, uint32_t cy = 0u, uint32_t cz = 0u) {
        ndep++;
        const bool hotbox = TALLY != MCX_TALLY_F32_SHADOW && kp.hot_cells;
        uint32_t x = cx, y = cy, z = cz;
        if (hotbox || (BLOCKED && kp.tblock)) {
            if (!(cx < kp.nx && cy < kp.ny && cz < kp.nz && cz * kp.dimxy + cy * kp.nx + cx == vox))
                vox_xyz(kp, vox, x, y, z);
        }
        if (hotbox && (uint32_t)tshift < kp.hot_g) {
            const uint32_t lx = x - kp.hot_x0, ly = y - kp.hot_y0, lz = z - kp.hot_z0;
            if (lx < kp.hot_bx && ly < kp.hot_by && lz < kp.hot_bz) {
                extern __shared__ F4 lds[];
                const uint32_t c = (((uint32_t)tshift * kp.hot_bz + lz) * kp.hot_by + ly) * kp.hot_bx + lx;
                if (TALLY == MCX_TALLY_F32_HILO) {
                    /* FP32 hi/lo cells: the returned old value gives the
                     * TwoSum error of the hi add, which goes to the lo half;
                     * chunks crossed leave the hi cell for the global chunk
                     * count, as in hilo_complete */
                    float* box = (float*)((char*)lds + hoff);
                    const float a = __hip_atomic_fetch_add(box + c, weight, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                    const float s = a + weight;
                    const float bb = s - a;
                    const float e = (a - (s - bb)) + (weight - bb);
                    if (e != 0.f) __hip_atomic_fetch_add(box + kp.hot_cells + c, e, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                    const float m = floorf(s * (1.f / MCX_HILO_CHUNK)) - floorf(a * (1.f / MCX_HILO_CHUNK));
                    if (m != 0.f) {
                        const float d = -m * MCX_HILO_CHUNK;
                        const float a2 = __hip_atomic_fetch_add(box + c, d, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                        const float s2 = a2 + d;
                        const float b2 = s2 - a2;
                        const float e2 = (a2 - (s2 - b2)) + (d - b2);
                        if (e2 != 0.f) __hip_atomic_fetch_add(box + kp.hot_cells + c, e2, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                        atom_add_nr(kp.fieldk + (size_t)vox + (size_t)tshift * kp.nvox, (int32_t)m);
                        const float x = __hip_atomic_exchange(box + kp.hot_cells + c, 0.f, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                        if (x != 0.f) {
                            const float a3 = __hip_atomic_fetch_add(box + c, x, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                            const float s3 = a3 + x;
                            const float b3 = s3 - a3;
                            const float e3 = (a3 - (s3 - b3)) + (x - b3);
                            if (e3 != 0.f) __hip_atomic_fetch_add(box + kp.hot_cells + c, e3, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                        }
                    }
                } else {
                    double* box = (double*)((char*)lds + hoff);
                    __hip_atomic_fetch_add(box + c, (double)weight, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                }
                nhot++;
                return;
            }
        }
        if (BLOCKED && kp.tblock) {
            const size_t c = block_cell(kp, x, y, z, (uint32_t)tshift);
            if (main) {
                if ((uint32_t)(c >> 3) == bopen) {
                    blk[c & 7u] += weight;
                    return;
                }
                if (pend_i == ~(size_t)0) {
                    pend_i = c;
                    pend_w = weight;
                    return;
                }
            }
            atom_add_nr(tally64 + c, (double)weight);
            nreq++;
            return;
        }
        nreq++;
        const size_t i = (size_t)vox + (size_t)tshift * kp.nvox;
        if (TALLY == MCX_TALLY_F64) {
            atom_add_nr(tally64 + i, (double)weight);  /* no return: fire and forget */
        } else if (TALLY == MCX_TALLY_F64_XCD) {
            /* only this XCD's workgroups touch this copy: the add can
             * complete in the XCD's own L2 */
            __hip_atomic_fetch_add(glb(tally64) + i, (double)weight, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        } else {
            flush_pending();
            pend_old = atom_add_ret(kp.field + i, weight);
            pend_w = weight;
            pend_i = i;
        }
    }

    /* work counters of this lane into kp.counters (when requested) */
    __device__ __forceinline__ void add_counters() const {
        if (kp.counters) {
            atomicAdd((unsigned long long*)kp.counters + 0, (unsigned long long)nstep);
            atomicAdd((unsigned long long*)kp.counters + 1, (unsigned long long)ndep);
            atomicAdd((unsigned long long*)kp.counters + 2, (unsigned long long)nscat);
            atomicAdd((unsigned long long*)kp.counters + 3, (unsigned long long)nlaunch);
            if (kp.hot_cells) atomicAdd((unsigned long long*)kp.counters + 4, (unsigned long long)nhot);
            atomicAdd((unsigned long long*)kp.counters + 5, (unsigned long long)nreq);
            /* the wave loops until its last lane is done: nstep + 1
             * iterations for its longest-stepping lane */
            const uint32_t wit = __reduce_max_sync(__activemask(), nstep) + 1u;
            if (__lane_id() == (uint32_t)__builtin_amdgcn_readfirstlane(__lane_id()))
                atomicAdd((unsigned long long*)kp.counters + 6, (unsigned long long)wit);
        }
    }

    /* src/mcx_core.cu:1767-2528, one lane */
    __device__ __forceinline__ void run() {
        const uint32_t flag = kp.savedetflag;
        const uint32_t NSCAT = (flag >> 1) & 1, PPATH = (flag >> 2) & 1, MOM = (flag >> 3) & 1;
        p = {0.f, 0.f, 0.f, __builtin_nanf("")};
        v = {0.f, 0.f, 0.f, 0.f};
        f = {0.f, 0.f, 0.f, -1.f};
        fd0 = fd1 = fd2 = 0;
        fd3 = -1;
        mediaid = mcx_f2bits(kp.srcparam2[3]);
        idx1d = 0;
        esc = glb(kp.genergy)[2 * idx];
        launched = glb(kp.genergy)[2 * idx + 1];
        winit = 0.f;
        pend_i = ~(size_t)0;
        tally64 = kp.field64;
        if (TALLY == MCX_TALLY_F64_XCD) tally64 += (size_t)xcc_id() * kp.fieldlen;
        if (TALLY == MCX_TALLY_F64 && kp.nmaps_mask) tally64 += (size_t)(idx & kp.nmaps_mask) * kp.mapstride;
        bopen = MCX_NO_BLOCK;
        if (BLOCKED && kp.tblock) {
            extern __shared__ F4 lds[];
            blk = (float*)((char*)lds + kp.blk_off) + 9u * threadIdx.x;
            for (uint32_t c = 0; c < 8; c++) blk[c] = 0.f;
            blk[8] = mcx_bits2f(MCX_NO_BLOCK);
        }
        nstep = ndep = nscat = nlaunch = nhot = nreq = 0;
        svsplit = svlo = svup = svupper = 0;
        rpx = rpy = rpz = nvx = nvy = nvz = 0.f;
        testint = hitintf = false;
        uint32_t idx1dold, mediaidold = 0, isdet = 0;
        float n1, len, slen;
        prop = tab[1];
        const MCX_GLB uint32_t* sd = glb(kp.seeds) + 4 * idx;
        t0 = ((uint64_t)sd[0] << 32) | sd[1];
        t1 = ((uint64_t)sd[2] << 32) | sd[3];
        /* Every termination sets `relaunch` and restarts the loop, so the
         * (large) launch code exists once; the operations and their order
         * are those of the reference, where launchnewphoton is called in
         * place at each termination site. */
        bool first = true, relaunch = true;
        uint32_t detarg = 0;
        while (true) {
            block_flush();
            if (relaunch) {
                if (launch(detarg)) {
                    if (first) return;
                    break;
                }
                if (first) {
                    calc_rv();
                    first = false;
                }
                isdet = mediaid & DET_MASK;
                mediaid &= MED_MASK;
                relaunch = false;
                testint = true;
            }
            if (!(f.w < (float)nphoton)) break;
            nstep++;
            if (++pstep > MAX_PHOTON_STEPS) {
                atom_add_u32(kp.detcount + 1, 1u);
                flush_pending();
                return;
            }
            if (f.x <= 0.f) {
                f.x = scatlen();
                if (v.w != EPSF) {
                    float cphi = 1.f, sphi = 0.f, stheta, ctheta, tmp0;
                    if (!kp.is2d) mcx_sincos2pi(rnd(), &sphi, &cphi);
                    tmp0 = (v.w > (float)kp.gscatter) ? 0.f : prop.z;
                    if (fabsf(tmp0) > EPSF) {
                        const float g = prop.z;
                        tmp0 = (1.f - g * g) / (1.f - g + 2.f * g * rnd());
                        tmp0 *= tmp0;
                        tmp0 = (1.f + g * g - tmp0) / (2.f * g);
                        tmp0 = fmaxf(-1.f, fminf(1.f, tmp0));
                        ctheta = tmp0;
                        stheta = mcx_sin_from_cos(ctheta);
                    } else {
                        ctheta = 2.f * rnd() - 1.f;
                        stheta = mcx_sin_from_cos(ctheta);
                    }
                    if (kp.savedet) {
                        const uint32_t lab = SPLIT ? sv_label() : (mediaid & MED_MASK);
                        if (!SPLIT || lab > 0) {
                            if (NSCAT) ppath(lab - 1) += 1.f;
                            if (MOM) ppath(kp.maxmedia * (NSCAT + PPATH) + lab - 1) += 1.f - ctheta;
                        }
                    }
                    if (kp.is2d)
                        rotatevector2d((rnd() > 0.5f ? stheta : -stheta), ctheta);
                    else
                        rotatevector(stheta, ctheta, sphi, cphi);
                    v.w += 1.f;
                    calc_rv();
                    nscat++;
                    /* src/mcx_core.cu:2019-2062: replayed photons tally their
                     * detected weight at every scattering event (otRFmus:
                     * times the modulation phasor of their time of flight,
                     * real and imaginary planes) */
                    if (kp.replay && (kp.outputtype == otWP || kp.outputtype == otDCS || kp.outputtype == otWPTOF ||
                                      kp.outputtype == otRFmus)) {
                        const float tof = glb(kp.replaytof)[gid];
                        int ts = (int)floorf((tof - kp.twin0) * kp.Rtstep);
                        ts = max(0, min((int)kp.maxgate - 1, ts));
                        if (kp.replaydet == -1) ts += ((glb(kp.replaydetid)[gid] & 0xFFFF) - 1) * (int)kp.maxgate;
                        if (kp.outputtype == otRFmus) {
                            float rs, rc;
                            mcx_sincosf(kp.omega * tof, &rs, &rc);
                            const float rw = glb(kp.replayweight)[gid];
                            deposit(idx1d, ts, rw * rc);
                            if (kp.omega > 0.f) deposit(idx1d, ts + kp.rfshift, rw * rs);
                        } else {
                            float w = (kp.outputtype == otDCS) ? (1.f - ctheta) : 1.f;
                            w = (kp.outputtype == otWPTOF) ? tof : w;
                            w *= glb(kp.replayweight)[gid];
                            deposit(idx1d, ts, w);
                        }
                    }
                }
                v.w = (float)(int)v.w;
                testint = true;
            }
            n1 = prop.w;
            /* in a split voxel the side's properties are kept */
            if (!SPLIT || !svsplit) setprop(mediaid, idx1d);
            len = hitgrid(p, fd0, fd1, fd2, fd3, rv0, rv1, rv2);
            const bool gs = (v.w + 1.f > (float)kp.gscatter);
            slen = len * prop.y * (gs ? (1.f - prop.z) : 1.f);
            slen = fminf(slen, f.x);
            len = ((prop.y == 0.f) ? len : (slen / (prop.y * (gs ? (1.f - prop.z) : 1.f))));
            if (SPLIT) hitintf = (svsplit && testint) ? ray_plane(len, slen) : false;
            p.x = p.x + len * v.x; p.y = p.y + len * v.y; p.z = p.z + len * v.z;
            const bool stop = (slen == f.x) || (SPLIT && hitintf);
            const uint32_t ox = (uint32_t)fd0, oy = (uint32_t)fd1, oz = (uint32_t)fd2;
            if (fd3 == 0) fd0 += stop ? 0 : (v.x > 0.f ? 1 : -1);
            if (fd3 == 1) fd1 += stop ? 0 : (v.y > 0.f ? 1 : -1);
            if (fd3 == 2) fd2 += stop ? 0 : (v.z > 0.f ? 1 : -1);
            p.w *= mcx_expf(-prop.x * len);
            f.x -= slen;
            f.y += len * prop.w * kp.oneoverc0;
            f.z += len;
            if (kp.savedet && PPATH) {
                const uint32_t lab = SPLIT ? sv_label() : (mediaid & MED_MASK);
                if (!SPLIT || lab > 0) ppath(kp.maxmedia * NSCAT + lab - 1) += len;
            }
            mediaidold = mediaid | (isdet & DET_MASK);
            idx1dold = idx1d;
            const bool inside = (uint32_t)fd0 < kp.nx && (uint32_t)fd1 < kp.ny && (uint32_t)fd2 < kp.nz;
            const uint32_t vox = (uint32_t)(fd2 * (int)kp.dimxy + fd1 * (int)kp.nx + fd0);
            /* The label of the voxel entered is loaded (at index 0 when the
             * photon left the grid) before the deposit for the voxel left,
             * which does not need it: the wait for this load then leaves
             * the tally atomic in flight until the next step's load. */
            const uint32_t label = load_media<MT>(kp.media, inside ? vox : 0u);
            idx1d = inside ? vox : ((fd0 < 0 || fd1 < 0 || fd2 < 0) ? OUT_MIN : OUT_MAX);
            const bool moved = ((idx1d != idx1dold) || (SPLIT && hitintf)) && mediaidold;
            if (moved) {
                if (kp.save2pt && f.y >= kp.twin0 && f.y < kp.twin1) {
                    float weight = 0.f, imag = 0.f;
                    int tshift = (int)floorf((f.y - kp.twin0) * kp.Rtstep);
                    if (tshift > (int)kp.maxgate - 1) tshift = (int)kp.maxgate - 1;
                    if (kp.outputtype == otEnergy)
                        weight = w0 - p.w;
                    else if (kp.outputtype == otFluence || kp.outputtype == otFlux)
                        weight = (prop.x < EPSF) ? (w0 * f.z) : ((w0 - p.w) / prop.x);
                    else if (kp.replay && (kp.outputtype == otJacobian || kp.outputtype == otRF || kp.outputtype == otWLTOF)) {
                        /* src/mcx_core.cu:2208-2220 (the gate index is clamped
                         * to the gate range, as the scattering tally clamps it
                         * to the last gate) */
                        const float tof = glb(kp.replaytof)[gid];
                        weight = glb(kp.replayweight)[gid] * f.z;
                        if (kp.outputtype == otRF) {
                            /* src/mcx_core.cu:2211-2213, :2254-2257, :2271-2273:
                             * real part -w L cos(omega tof), imaginary part
                             * -w L sin(omega tof), or p.w L sin(omega tof)
                             * for pattern sources */
                            float rs, rc;
                            mcx_sincosf(kp.omega * tof, &rs, &rc);
                            imag = (kp.srctype == 5 || kp.srctype == 15) ? p.w * f.z * rs : -weight * rs;
                            weight = -weight * rc;
                        } else if (kp.outputtype == otWLTOF) {
                            weight = weight * tof;
                        }
                        tshift = max(0, min((int)kp.maxgate - 1, (int)floorf((tof - kp.twin0) * kp.Rtstep)));
                        if (kp.replaydet == -1) tshift += ((glb(kp.replaydetid)[gid] & 0xFFFF) - 1) * (int)kp.maxgate;
                    } else if (kp.outputtype == otL)
                        weight = w0 * f.z;
                    if (fabsf(weight) > 0.f) deposit(idx1dold, tshift, weight, true, ox, oy, oz);
                    if (kp.rfshift && kp.omega > 0.f && fabsf(imag) > 0.f)
                        deposit(idx1dold, tshift + kp.rfshift, imag, false, ox, oy, oz);
                }
                w0 = p.w;
                f.z = 0.f;
            }
            if (inside) {
                isdet = label & DET_MASK;
                mediaid = label & MED_MASK;
            } else {
                mediaid = 0;
                isdet = (kp.bcface >> (4 * ((idx1d == OUT_MAX) * 3 + fd3))) & 0xFu;
            }
            if (!moved) mediaid = mediaidold;
            /* src/mcx_core.cu:2294-2304: a new voxel gets its split state; a
             * crossing of the interface moves the photon to the other side */
            if (SPLIT) {
                if (idx1d != idx1dold) {
                    setprop(mediaid, idx1d);
                    testint = true;
                } else if (hitintf) {
                    nvx = -nvx;
                    nvy = -nvy;
                    nvz = -nvz;
                    svupper = !svupper;
                    testint = false;
                }
            }
            const uint32_t bcf = isdet & 0xF;
            if ((mediaid == 0 &&
                 ((!ISREFLECT || (ISREFLECT && n1 == tab[0].w)) || (bcf == bcUnknown && !kp.doreflect) ||
                  bcf == bcAbsorb || bcf == bcCyclic) &&
                 bcf != bcMirror) ||
                (SPLIT && (idx1d != idx1dold || hitintf) && !svupper && !svlo && (!ISREFLECT || (ISREFLECT && n1 == tab[0].w))) ||
                f.y > kp.twin1) {
                if (isdet == bcCyclic) {
                    if (fd3 == 0) {
                        p.x = nextafter_mcx(roundf(p.x + ((idx1d == OUT_MIN) ? kp.maxidx_x : -kp.maxidx_x)), (v.x > 0.f) - (v.x < 0.f), kp.maxvoidstep);
                        fd0 = (int)floorf(p.x);
                    }
                    if (fd3 == 1) {
                        p.y = nextafter_mcx(roundf(p.y + ((idx1d == OUT_MIN) ? kp.maxidx_y : -kp.maxidx_y)), (v.y > 0.f) - (v.y < 0.f), kp.maxvoidstep);
                        fd1 = (int)floorf(p.y);
                    }
                    if (fd3 == 2) {
                        p.z = nextafter_mcx(roundf(p.z + ((idx1d == OUT_MIN) ? kp.maxidx_z : -kp.maxidx_z)), (v.z > 0.f) - (v.z < 0.f), kp.maxvoidstep);
                        fd2 = (int)floorf(p.z);
                    }
                    if ((uint32_t)fd0 < kp.nx && (uint32_t)fd1 < kp.ny && (uint32_t)fd2 < kp.nz) {
                        idx1d = (uint32_t)(fd2 * (int)kp.dimxy + fd1 * (int)kp.nx + fd0);
                        mediaid = load_media<MT>(kp.media, idx1d);
                        isdet = mediaid & DET_MASK;
                        mediaid &= MED_MASK;
                        continue;
                    }
                }
                detarg = (((idx1d == OUT_MAX || idx1d == OUT_MIN) && ((kp.bcdet >> ((idx1d == OUT_MAX) * 3 + fd3)) & 1u))
                              ? OUT_MIN : (mediaidold & DET_MASK));
                relaunch = true;
                continue;
            }
            if (fabsf(p.w) < kp.minenergy) {
                if (rnd() * ROULETTE <= 1.f) {
                    p.w *= ROULETTE;
                } else {
                    detarg = mediaidold & DET_MASK;
                    relaunch = true;
                    continue;
                }
            }
            if (ISREFLECT && SPLIT && hitintf) {
                /* src/mcx_core.cu:2386-2404: at the interface plane */
                if (tab[svlo].w != tab[svup].w) {
                    nvx = -nvx;
                    nvy = -nvy;
                    nvz = -nvz;
                    if (reflectray(n1)) {
                        detarg = mediaidold & DET_MASK;
                        relaunch = true;
                        continue;
                    }
                } else {
                    prop = tab[sv_label()];
                }
            } else if (ISREFLECT) {
                /* src/mcx_core.cu:2381-2416: formats from 100 on keep one
                 * refractive index (row 1) and decode the entered voxel only
                 * when a reflection test follows; split voxels were decoded
                 * on entry */
                const bool lt100 = !WORDS || kp.mediaformat < 100;
                if (lt100 && !SPLIT) setprop(mediaid, idx1d);
                if (((mediaid && kp.doreflect) ||
                     (mediaid == 0 && ((bcf == bcUnknown && kp.doreflect) || bcf == bcReflect || bcf == bcMirror))) &&
                    (bcf == bcMirror || n1 != (lt100 ? prop.w : tab[mediaid > 0u ? 1 : 0].w))) {
                    float Rtotal = 1.f;
                    float cphi, sphi, stheta, ctheta, tmp0, tmp1;
                    if (WORDS && !SPLIT) setprop(mediaid, idx1d);
                    tmp0 = n1 * n1;
                    tmp1 = prop.w * prop.w;
                    cphi = fabsf(vcomp(fd3));
                    sphi = 1.f - cphi * cphi;
                    len = 1.f - tmp0 / tmp1 * sphi;
                    if (len > 0.f && bcf != bcMirror) {
                        ctheta = tmp0 * cphi * cphi + tmp1 * len;
                        stheta = 2.f * n1 * prop.w * cphi * __builtin_sqrtf(len);
                        Rtotal = (ctheta - stheta) / (ctheta + stheta);
                        ctheta = tmp1 * cphi * cphi + tmp0 * len;
                        Rtotal = (Rtotal + (ctheta - stheta) / (ctheta + stheta)) * 0.5f;
                    }
                    if (Rtotal < 1.f && (!(mediaid == 0 && bcf == bcMirror)) && rnd() > Rtotal) {
                        transmit(n1, prop.w, fd3);
                        if (mediaid == 0 || (SPLIT && sv_label() == 0)) {
                            detarg = (((idx1d == OUT_MAX || idx1d == OUT_MIN) && ((kp.bcdet >> ((idx1d == OUT_MAX) * 3 + fd3)) & 1u))
                                          ? OUT_MIN : (mediaidold & DET_MASK));
                            relaunch = true;
                            continue;
                        }
                        calc_rv();
                    } else {
                        if (fd3 == 0) {
                            v.x = -v.x;
                        } else if (fd3 == 1) {
                            v.y = -v.y;
                        } else {
                            v.z = -v.z;
                        }
                        calc_rv();
                        if (fd3 == 0) {
                            p.x = nextafter_mcx(__builtin_rintf(p.x), (v.x > 0.f) - (v.x < 0.f), kp.maxvoidstep);
                            fd0 = (int)floorf(p.x);
                        } else if (fd3 == 1) {
                            p.y = nextafter_mcx(__builtin_rintf(p.y), (v.y > 0.f) - (v.y < 0.f), kp.maxvoidstep);
                            fd1 = (int)floorf(p.y);
                        } else {
                            p.z = nextafter_mcx(__builtin_rintf(p.z), (v.z > 0.f) - (v.z < 0.f), kp.maxvoidstep);
                            fd2 = (int)floorf(p.z);
                        }
                        idx1d = idx1dold;
                        mediaid = load_media<MT>(kp.media, idx1d) & MED_MASK;
                        setprop(mediaid, idx1d);
                        if (SPLIT && sv_label() == 0) {
                            detarg = mediaidold & DET_MASK;
                            relaunch = true;
                            continue;
                        }
                        n1 = prop.w;
                    }
                } else if (lt100 && !SPLIT) {
                    setprop(mediaidold, idx1d);
                }
            } else if (SPLIT) {
                prop = tab[sv_label()];
            }
            if (mediaid == 0 || idx1d == OUT_MIN || idx1d == OUT_MAX) {
                atom_add_u32(kp.detcount + 1, 1u);
                flush_pending();
                return;
            }
        }
        flush_pending();
        glb(kp.genergy)[2 * idx] = esc;
        glb(kp.genergy)[2 * idx + 1] = launched;
    }
};

}  // namespace mcxk

using namespace mcxk;

/* Hot-cell box (KParam::hot_*) in LDS at byte offset `off`: FP64 cells for
 * the FP64 tallies, FP32 hi then lo halves for the hi/lo tally (8 bytes per
 * cell either way).  Both helpers are run by `n` lanes with indices
 * `t` = 0..n-1 of one workgroup (single simulation) or one wave (batch). */
template <int TALLY>
__device__ __forceinline__ void hot_box_zero(const KParam& kp, uint32_t off, uint32_t t, uint32_t n) {
    extern __shared__ F4 lds[];
    uint2* box = (uint2*)((char*)lds + off);
    for (uint32_t c = t; c < kp.hot_cells; c += n) box[c] = uint2{0u, 0u};
}

/* adds the box to the tally: one global atomic per non-zero cell (hi/lo:
 * the hi half by a returning add completed by hilo_complete) */
template <int TALLY>
__device__ __forceinline__ void hot_box_flush(const KParam& kp, uint32_t off, uint32_t t, uint32_t n) {
    extern __shared__ F4 lds[];
    const char* base = (const char*)lds + off;
    double* t64 = kp.field64;
    if (TALLY == MCX_TALLY_F64_XCD) t64 += (size_t)xcc_id() * kp.fieldlen;
    const uint32_t plane = kp.hot_bx * kp.hot_by, vol = plane * kp.hot_bz;
    for (uint32_t c = t; c < kp.hot_cells; c += n) {
        const uint32_t g = c / vol, r = c - g * vol, lz = r / plane, r2 = r - lz * plane, ly = r2 / kp.hot_bx;
        const uint32_t lx = r2 - ly * kp.hot_bx;
        const size_t i = (TALLY == MCX_TALLY_F64 && kp.tblock)
                             ? block_cell(kp, kp.hot_x0 + lx, kp.hot_y0 + ly, kp.hot_z0 + lz, g)
                             : (size_t)((kp.hot_z0 + lz) * kp.dimxy + (kp.hot_y0 + ly) * kp.nx + kp.hot_x0 + lx) + (size_t)g * kp.nvox;
        if (TALLY == MCX_TALLY_F32_HILO) {
            const float w = ((const float*)base)[c];
            const float lo = ((const float*)base)[kp.hot_cells + c];
            if (w != 0.f) hilo_complete(kp.field + i, kp.field2 + i, kp.fieldk + i, atom_add_ret(kp.field + i, w), w);
            if (lo != 0.f) atom_add_nr(kp.field2 + i, lo);
        } else {
            const double sum = ((const double*)base)[c];
            if (sum == 0.0) continue;
            if (TALLY == MCX_TALLY_F64)
                atom_add_nr(t64 + i, sum);
            else
                __hip_atomic_fetch_add(glb(t64) + i, sum, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        }
    }
}

/* The parameter block is read through a device pointer: every field is a
 * wave-uniform load served by the scalar cache.  (Binding a by-value kernel
 * argument to a reference would copy the whole block to scratch.) */
#ifndef MCX_MIN_WAVES_PER_SIMD
#define MCX_MIN_WAVES_PER_SIMD 1
#endif
/* Label media: up to 1024 lanes a workgroup, so that with the hot-cell box
 * the host may launch one workgroup per CU holding all of its waves and the
 * box takes the CU's whole LDS (these variants fit 128 VGPRs without
 * spilling).  The property-word and split-voxel variants need 126-139 VGPRs
 * and keep 256-lane workgroups rather than spill to fit 4 waves per SIMD. */
template <typename MT>
struct TransportLanes {
    static constexpr int value = 1024;
};
template <>
struct TransportLanes<mcxk::MediaWord> {
    static constexpr int value = 256;
};
template <>
struct TransportLanes<mcxk::MediaSplit> {
    static constexpr int value = 256;
};
template <int ISPENCIL, int ISREFLECT, int TALLY, typename MT>
__global__ void __launch_bounds__(TransportLanes<MT>::value, MCX_MIN_WAVES_PER_SIMD)
    mcx_transport(const KParam* __restrict__ kpp) {
    const KParam& kp = *kpp;
    extern __shared__ F4 lds[];
    for (uint32_t i = threadIdx.x; i < kp.ntable; i += blockDim.x) {
        const MCX_GLB float* g = glb(kp.gprop) + 4 * i;
        lds[i] = F4{g[0], g[1], g[2], g[3]};
    }
    const bool hot = TALLY != MCX_TALLY_F32_SHADOW && kp.hot_cells;
    if (hot) hot_box_zero<TALLY>(kp, kp.hot_off, threadIdx.x, blockDim.x);
    __syncthreads();
    const uint32_t idx = blockIdx.x * blockDim.x + threadIdx.x;
    const long long total = (long long)kp.threadphoton * kp.nthread + kp.oddphotons;
    if ((long long)idx < total && idx < kp.nthread) {
        float* ppbase = (float*)(lds + kp.ntable) + threadIdx.x;
        Lane<ISPENCIL, ISREFLECT, TALLY, MT> L(kp, lds, ppbase, blockDim.x, idx);
        L.nphoton = kp.threadphoton + ((int)idx < kp.oddphotons);
        for (uint32_t i = 0; i < kp.partialdata; i++) L.ppath(i) = 0.f;
        L.run();
        L.add_counters();
    }
    if (!hot) return;
    /* every lane of the workgroup has finished: add the box to the tally */
    __syncthreads();
    hot_box_flush<TALLY>(kp, kp.hot_off, threadIdx.x, blockDim.x);
}

/* Batched persistent launch: every wave repeatedly takes the next chunk of
 * 64 virtual threads from a global queue; a chunk belongs to one instance,
 * so the instance parameters stay wave-uniform (scalar loads) and each
 * wave keeps that instance's property table in its own LDS slice.  Chunks
 * are numbered in `order` (most expensive instances first) with
 * chunk_start[k] the first chunk of the k-th instance in that order.  A
 * virtual thread's work is identical to the single-instance kernel's, so a
 * batched instance reproduces its plain run lane for lane. */
/* Held to 128 VGPRs (4 waves per SIMD): the hi/lo variant otherwise takes
 * 141 and runs 3 waves per SIMD; the 24 B per lane it then spills cost
 * less than the fourth wave gains (batch256 at quarter scale: 5281 against
 * 5941 ms per graph). */
template <int TALLY>
__global__ void __attribute__((amdgpu_waves_per_eu(4))) __launch_bounds__(256, MCX_MIN_WAVES_PER_SIMD)
mcx_transport_batch(const KParam* __restrict__ params, const uint32_t* __restrict__ chunk_start,
                    const uint32_t* __restrict__ order, uint32_t ninst, uint32_t total_chunks, uint32_t maxtab,
                    uint32_t boxcells, uint32_t* __restrict__ queue) {
    extern __shared__ F4 lds[];
    const uint32_t wave = threadIdx.x >> 6, lane = threadIdx.x & 63u;
    /* per wave: the property table, then the hot-cell box (boxcells x 8 B) */
    const uint32_t wstride = maxtab * 16u + boxcells * 8u;
    F4* tab = (F4*)((char*)lds + wave * wstride);
    const uint32_t boff = wave * wstride + maxtab * 16u;
    while (true) {
        uint32_t c = 0;
        if (lane == 0) c = atomicAdd(queue, 1u);
        c = __builtin_amdgcn_readfirstlane(__shfl(c, 0));
        if (c >= total_chunks) break;
        uint32_t lo = 0, hi = ninst;          /* last k with chunk_start[k] <= c */
        while (hi - lo > 1) {
            const uint32_t mid = (lo + hi) >> 1;
            if (chunk_start[mid] <= c) lo = mid; else hi = mid;
        }
        const uint32_t inst = __builtin_amdgcn_readfirstlane(order[lo]);
        const KParam& kp = params[inst];
        for (uint32_t i = lane; i < kp.ntable; i += 64) {
            const MCX_GLB float* g = glb(kp.gprop) + 4 * i;
            tab[i] = F4{g[0], g[1], g[2], g[3]};
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        const bool hot = kp.hot_cells != 0;
        if (hot) hot_box_zero<TALLY>(kp, boff, lane, 64);
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        const uint32_t idx = (c - chunk_start[lo]) * 64u + lane;
        const long long total = (long long)kp.threadphoton * kp.nthread + kp.oddphotons;
        if (idx < kp.nthread && (long long)idx < total) {
            Lane<2, 1, TALLY, uint8_t> L(kp, tab, (float*)tab, 64, idx);
            L.hoff = boff;
            L.nphoton = kp.threadphoton + ((int)idx < kp.oddphotons);
            L.run();
            L.add_counters();
        }
        /* the chunk's lanes are done: add the wave's box to the instance's tally */
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        if (hot) hot_box_flush<TALLY>(kp, boff, lane, 64);
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }
}

/* FP64 sums of the per-lane escaped / launched energies of each instance */
__global__ void mcx_batch_energy(const KParam* __restrict__ params, BatchInst* __restrict__ bi) {
    __shared__ double red[2][256];
    const KParam& kp = params[blockIdx.x];
    double e = 0.0, t = 0.0;
    for (uint32_t i = threadIdx.x; i < kp.nthread; i += blockDim.x) {
        e += (double)kp.genergy[2 * i];
        t += (double)kp.genergy[2 * i + 1];
    }
    red[0][threadIdx.x] = e;
    red[1][threadIdx.x] = t;
    __syncthreads();
    for (uint32_t s = blockDim.x / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s) {
            red[0][threadIdx.x] += red[0][threadIdx.x + s];
            red[1][threadIdx.x] += red[1][threadIdx.x + s];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        bi[blockIdx.x].energy[0] = red[0][0];
        bi[blockIdx.x].energy[1] = red[1][0];
    }
}

/* merge the tally halves in FP64 and normalise, one row of blocks per
 * instance (blockIdx.y).  The scale is the single-simulation path's:
 * unitinmm / (energytot * Vvox * tstep) in FP64, rounded to FP32 (times
 * tstep for fluence), applied to the FP32-rounded merged value. */
template <int TALLY>
__global__ void mcx_batch_merge(const KParam* __restrict__ params, const BatchInst* __restrict__ bi) {
    const KParam& kp = params[blockIdx.y];
    const BatchInst& b = bi[blockIdx.y];
    const double et = b.energy[1];
    float scale = 1.f;
    if (b.normalize) {
        if (b.outputtype == otFlux || b.outputtype == otFluence) {
            scale = (float)((double)b.unitinmm / (et * (double)b.vvox * (double)b.tstep));
            if (b.outputtype == otFluence) scale *= b.tstep;
        } else if (b.outputtype == otEnergy || b.outputtype == otL) {
            scale = (float)(1.0 / et);
        }
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) b.energy[2] = (double)scale;
    const size_t n = kp.fieldlen;
    MCX_GLB float* out = glb(b.out);
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        float v;
        if (TALLY == MCX_TALLY_F64)
            v = (float)glb(kp.field64)[i];
        else
            v = (float)((double)glb(kp.fieldk)[i] * (double)MCX_HILO_CHUNK + (double)glb(kp.field)[i] + (double)glb(kp.field2)[i]);
        out[i] = v * scale;
    }
}

/* time-integrated (CW) map of each instance's normalised output, one row
 * of blocks per instance: cw[v] = tstep * sum over gates of out[g][v],
 * summed in FP64 in gate order (one thread per voxel, gate planes read
 * coalesced) */
__global__ void mcx_batch_cw(const KParam* __restrict__ params, const BatchInst* __restrict__ bi) {
    const KParam& kp = params[blockIdx.y];
    const BatchInst& b = bi[blockIdx.y];
    const size_t nvox = kp.nvox;
    const uint32_t gates = kp.maxgate;
    const MCX_GLB float* out = glb(b.out);
    MCX_GLB float* cw = glb(b.cw);
    for (size_t v = (size_t)blockIdx.x * blockDim.x + threadIdx.x; v < nvox; v += (size_t)gridDim.x * blockDim.x) {
        double acc = 0.0;
        for (uint32_t g = 0; g < gates; g++) acc += (double)out[(size_t)g * nvox + v];
        cw[v] = (float)(acc * (double)b.tstep);
    }
}

/* zero fill of 16-byte words (arena clears inside a captured graph) */
__global__ void mcx_fill_zero16(uint4* __restrict__ p, size_t n) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        p[i] = make_uint4(0u, 0u, 0u, 0u);
}

__global__ void mcx_fill_u32(uint32_t* __restrict__ p, uint32_t v, size_t n) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) p[i] = v;
}

/* src/mcx_core.cu:1731-1742: n draws of one lane's stream */
__global__ void mcx_test_rng(float* out, const uint32_t* seeds, uint32_t n) {
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    uint64_t t0 = ((uint64_t)seeds[0] << 32) | seeds[1];
    uint64_t t1 = ((uint64_t)seeds[2] << 32) | seeds[3];
    for (uint32_t i = 0; i < n; i++) {
        uint64_t s1 = t0;
        const uint64_t s0 = t1;
        t0 = s0;
        s1 ^= s1 << 23;
        t1 = s1 ^ s0 ^ (s1 >> 18) ^ (s0 >> 5);
        uint32_t lo = (uint32_t)(t1 + s0);
        out[i] = mcx_bits2f(0x3F800000u | (lo >> 9)) - 1.0f;
    }
}

/* Tally merge, vectorised: out = raw + shadow (mode 0, FP32 as the
 * reference's host loop src/mcx_core.cu:3687-3696), or the FP64 sum of the
 * hi/lo pair rounded once to FP32 (mode 1), or FP64 rounded (mode 2). */
__global__ void mcx_merge_f32(float* __restrict__ out, const float* __restrict__ a, const float* __restrict__ b, size_t n) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) out[i] = a[i] + b[i];
}

__global__ void mcx_merge_hilo(float* __restrict__ out, const float* __restrict__ hi, const float* __restrict__ lo,
                               const int32_t* __restrict__ chunks, size_t n) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) out[i] = (float)((double)chunks[i] * (double)MCX_HILO_CHUNK + (double)hi[i] + (double)lo[i]);
}

__global__ void mcx_merge_f64x(float* __restrict__ out, const double* __restrict__ d, size_t n) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) {
        double t = 0.0;
        for (int c = 0; c < MCX_NUM_XCD; c++) t += d[(size_t)c * n + i];
        out[i] = (float)t;
    }
}

__global__ void mcx_merge_f64(float* __restrict__ out, const double* __restrict__ d, size_t n) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) out[i] = (float)d[i];
}

/* blocked FP64 tally (KParam::tblock) -> FP32 volume in the reference's
 * [gate][z][y][x] order; one lane per output cell */
__global__ void mcx_merge_f64_blocked(float* __restrict__ out, const double* __restrict__ d, uint32_t nx, uint32_t ny,
                                      uint32_t nz, uint32_t bnx, uint32_t bny, uint32_t bnz, size_t n) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    const size_t nvox = (size_t)nx * ny * nz;
    for (; i < n; i += stride) {
        const uint32_t g = (uint32_t)(i / nvox);
        const uint32_t r = (uint32_t)(i - (size_t)g * nvox);
        const uint32_t z = r / (nx * ny), r2 = r - z * nx * ny, y = r2 / nx, x = r2 - y * nx;
        const size_t b = (((size_t)g * bnz + (z >> 1)) * bny + (y >> 1)) * bnx + (x >> 1);
        out[i] = (float)d[(b << 3) + (x & 1u) + ((y & 1u) << 1) + ((z & 1u) << 2)];
    }
}

/* Replicas: the time-integrated map of replica m = tstep x the sum over
 * its gates of the normalised tally (scale[m] includes tstep), FP64 tally
 * linear or blocked (bnx > 0); one lane per (replica, voxel) */
__global__ void mcx_replica_cw(float* __restrict__ cw, const double* __restrict__ d, const double* __restrict__ scale,
                               uint32_t nx, uint32_t ny, uint32_t nz, uint32_t planes, uint32_t nmaps, uint32_t bnx,
                               uint32_t bny, uint32_t bnz) {
    const size_t nvox = (size_t)nx * ny * nz;
    const size_t n = nvox * nmaps;
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) {
        const uint32_t m = (uint32_t)(i / nvox);
        const uint32_t v = (uint32_t)(i - (size_t)m * nvox);
        const uint32_t z = v / (nx * ny), r = v - z * nx * ny, y = r / nx, x = r - y * nx;
        double t = 0.0;
        for (uint32_t g = 0; g < planes; g++) {
            const uint32_t gp = m * planes + g;
            if (bnx) {
                const size_t b = (((size_t)gp * bnz + (z >> 1)) * bny + (y >> 1)) * bnx + (x >> 1);
                t += d[(b << 3) + (x & 1u) + ((y & 1u) << 1) + ((z & 1u) << 2)];
            } else {
                t += d[(size_t)gp * nvox + v];
            }
        }
        cw[i] = (float)(t * scale[m]);
    }
}

extern "C" hipError_t mcxk_launch_replica_cw(float* cw, const double* d, const double* scale, uint32_t nx, uint32_t ny,
                                             uint32_t nz, uint32_t planes, uint32_t nmaps, uint32_t bnx, uint32_t bny,
                                             uint32_t bnz, hipStream_t stream) {
    const size_t n = (size_t)nx * ny * nz * nmaps;
    size_t g = (n + 255) / 256;
    const unsigned int grid = (unsigned int)(g > 65536 ? 65536 : (g ? g : 1));
    hipLaunchKernelGGL(mcx_replica_cw, dim3(grid), dim3(256), 0, stream, cw, d, scale, nx, ny, nz, planes, nmaps, bnx, bny,
                       bnz);
    return hipGetLastError();
}

/* ---- host-side dispatch (called from mcx_host.cpp) ------------------- */

typedef void (*mcx_kernel_fn)(const KParam*);

/* media kind: 0 uint8 labels, 1 uint32 labels, 2 non-label 32-bit words,
 * 3 split voxels */
template <int P, int R, int T>
static mcx_kernel_fn pick_mt(int media32) {
    if (media32 == 3) return (mcx_kernel_fn)mcx_transport<P, R, T, MediaSplit>;
    if (media32 == 2) return (mcx_kernel_fn)mcx_transport<P, R, T, MediaWord>;
    return media32 ? (mcx_kernel_fn)mcx_transport<P, R, T, uint32_t> : (mcx_kernel_fn)mcx_transport<P, R, T, uint8_t>;
}

template <int P, int R>
static mcx_kernel_fn pick_tally(int tally, int media32) {
    switch (tally) {
        case MCX_TALLY_F32_HILO: return pick_mt<P, R, MCX_TALLY_F32_HILO>(media32);
        case MCX_TALLY_F64: return pick_mt<P, R, MCX_TALLY_F64>(media32);
        case MCX_TALLY_F64_XCD: return pick_mt<P, R, MCX_TALLY_F64_XCD>(media32);
        default: return pick_mt<P, R, MCX_TALLY_F32_SHADOW>(media32);
    }
}

extern "C" hipError_t mcxk_launch_transport(const KParam* dkp, int ispencil, int isreflect, int tally, int media32,
                                            unsigned int grid, unsigned int block, size_t shmem, hipStream_t stream) {
    mcx_kernel_fn fn;
    if (ispencil)
        fn = isreflect ? pick_tally<1, 1>(tally, media32) : pick_tally<1, 0>(tally, media32);
    else
        fn = isreflect ? pick_tally<0, 1>(tally, media32) : pick_tally<0, 0>(tally, media32);
    if (shmem > 65536) {
        const hipError_t e = hipFuncSetAttribute((const void*)fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)shmem);
        if (e != hipSuccess) return e;
    }
    hipLaunchKernelGGL(fn, dim3(grid), dim3(block), shmem, stream, dkp);
    return hipGetLastError();
}

extern "C" hipError_t mcxk_kernel_attrs(int ispencil, int isreflect, int tally, int media32, int* numregs, int* maxthreads) {
    mcx_kernel_fn fn;
    if (ispencil)
        fn = isreflect ? pick_tally<1, 1>(tally, media32) : pick_tally<1, 0>(tally, media32);
    else
        fn = isreflect ? pick_tally<0, 1>(tally, media32) : pick_tally<0, 0>(tally, media32);
    hipFuncAttributes a;
    hipError_t e = hipFuncGetAttributes(&a, (const void*)fn);
    if (e == hipSuccess) {
        *numregs = a.numRegs;
        *maxthreads = a.maxThreadsPerBlock;
    }
    return e;
}

/* boxcells: hot-cell box cells per wave (0: none; every instance's
 * KParam::hot_cells is at most this) */
extern "C" hipError_t mcxk_launch_batch(int tally, const KParam* params, const uint32_t* chunk_start, const uint32_t* order,
                                        uint32_t ninst, uint32_t total_chunks, uint32_t maxtab, uint32_t boxcells,
                                        uint32_t* queue, unsigned int grid, unsigned int block, hipStream_t stream) {
    const size_t shmem = (size_t)(block / 64) * ((size_t)maxtab * sizeof(F4) + (size_t)boxcells * 8);
    if (tally == MCX_TALLY_F32_HILO)
        hipLaunchKernelGGL(mcx_transport_batch<MCX_TALLY_F32_HILO>, dim3(grid), dim3(block), shmem, stream, params, chunk_start,
                           order, ninst, total_chunks, maxtab, boxcells, queue);
    else if (tally == MCX_TALLY_F64)
        hipLaunchKernelGGL(mcx_transport_batch<MCX_TALLY_F64>, dim3(grid), dim3(block), shmem, stream, params, chunk_start,
                           order, ninst, total_chunks, maxtab, boxcells, queue);
    else
        return hipErrorInvalidValue;
    return hipGetLastError();
}

extern "C" hipError_t mcxk_batch_attrs(int tally, int* numregs) {
    hipFuncAttributes a;
    hipError_t e = hipFuncGetAttributes(&a, tally == MCX_TALLY_F64 ? (const void*)mcx_transport_batch<MCX_TALLY_F64>
                                                                   : (const void*)mcx_transport_batch<MCX_TALLY_F32_HILO>);
    if (e == hipSuccess) *numregs = a.numRegs;
    return e;
}

extern "C" hipError_t mcxk_launch_batch_post(int tally, const KParam* params, BatchInst* bi, uint32_t ninst, hipStream_t stream) {
    hipLaunchKernelGGL(mcx_batch_energy, dim3(ninst), dim3(256), 0, stream, params, bi);
    if (tally == MCX_TALLY_F64)
        hipLaunchKernelGGL(mcx_batch_merge<MCX_TALLY_F64>, dim3(512, ninst), dim3(256), 0, stream, params, (const BatchInst*)bi);
    else
        hipLaunchKernelGGL(mcx_batch_merge<MCX_TALLY_F32_HILO>, dim3(512, ninst), dim3(256), 0, stream, params, (const BatchInst*)bi);
    return hipGetLastError();
}

extern "C" hipError_t mcxk_launch_batch_cw(const KParam* params, const BatchInst* bi, uint32_t ninst, hipStream_t stream) {
    hipLaunchKernelGGL(mcx_batch_cw, dim3(256, ninst), dim3(256), 0, stream, params, bi);
    return hipGetLastError();
}

extern "C" hipError_t mcxk_fill_zero16(void* p, size_t bytes, hipStream_t stream) {
    const size_t n = bytes / 16;
    size_t g = (n + 255) / 256;
    hipLaunchKernelGGL(mcx_fill_zero16, dim3((unsigned int)(g > 8192 ? 8192 : (g ? g : 1))), dim3(256), 0, stream, (uint4*)p, n);
    return hipGetLastError();
}

extern "C" hipError_t mcxk_fill_u32(uint32_t* p, uint32_t v, size_t n, hipStream_t stream) {
    size_t g = (n + 255) / 256;
    hipLaunchKernelGGL(mcx_fill_u32, dim3((unsigned int)(g > 4096 ? 4096 : (g ? g : 1))), dim3(256), 0, stream, p, v, n);
    return hipGetLastError();
}

extern "C" hipError_t mcxk_launch_rngtest(float* out, const uint32_t* seeds, uint32_t n, hipStream_t stream) {
    hipLaunchKernelGGL(mcx_test_rng, dim3(1), dim3(64), 0, stream, out, seeds, n);
    return hipGetLastError();
}

extern "C" hipError_t mcxk_launch_merge_blocked(float* out, const double* a, uint32_t nx, uint32_t ny, uint32_t nz,
                                                 uint32_t bnx, uint32_t bny, uint32_t bnz, size_t n, hipStream_t stream) {
    const unsigned int block = 256;
    size_t g = (n + block - 1) / block;
    unsigned int grid = (unsigned int)(g > 65536 ? 65536 : (g ? g : 1));
    hipLaunchKernelGGL(mcx_merge_f64_blocked, dim3(grid), dim3(block), 0, stream, out, a, nx, ny, nz, bnx, bny, bnz, n);
    return hipGetLastError();
}

extern "C" hipError_t mcxk_launch_merge(int tally, float* out, const void* a, const void* b, const void* c, size_t n, hipStream_t stream) {
    const unsigned int block = 256;
    size_t g = (n + block - 1) / block;
    unsigned int grid = (unsigned int)(g > 65536 ? 65536 : (g ? g : 1));
    if (tally == MCX_TALLY_F32_HILO)
        hipLaunchKernelGGL(mcx_merge_hilo, dim3(grid), dim3(block), 0, stream, out, (const float*)a, (const float*)b,
                           (const int32_t*)c, n);
    else if (tally == MCX_TALLY_F64)
        hipLaunchKernelGGL(mcx_merge_f64, dim3(grid), dim3(block), 0, stream, out, (const double*)a, n);
    else if (tally == MCX_TALLY_F64_XCD)
        hipLaunchKernelGGL(mcx_merge_f64x, dim3(grid), dim3(block), 0, stream, out, (const double*)a, n);
    else
        hipLaunchKernelGGL(mcx_merge_f32, dim3(grid), dim3(block), 0, stream, out, (const float*)a, (const float*)b, n);
    return hipGetLastError();
}
