/*
 * mcx_config.cpp -- Config defaults, validation and preprocessing.
 *
 * Restates the reference's host-side input preparation for the transport
 * path:
 *   mcx_initcfg      src/mcx_utils.c:248-390
 *   mcx_clearcfg     src/mcx_utils.c:411-470
 *   mcx_preprocess   src/mcx_utils.c:1525-1820
 *   mcx_maskdet      src/mcx_utils.c:4214-4340
 *   mcx_validatecfg  src/mcx_utils.c:3748-3859
 *   mcx_normalize    src/mcx_utils.c:1227-1237
 *   mcx_error        src/mcx_utils.c:1314-1327
 */
#include <ctype.h>
#include <stddef.h>
#include <math.h>
#include <stdarg.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include <algorithm>
#include <mutex>
#include <vector>

#include "../../include/mcx_amd_ext.h"
#include "mcx_internal.h"

namespace mcx {

static thread_local int g_depth = 0;
static std::mutex g_errmu;
static int g_errid = 0;
static std::string g_errmsg;

EntryGuard::EntryGuard() { g_depth++; }
EntryGuard::~EntryGuard() { g_depth--; }

void record_error(int id, const char* msg) {
    std::lock_guard<std::mutex> lk(g_errmu);
    g_errid = id;
    g_errmsg = msg ? msg : "";
}

void handle_error(const Error& e, const Config* cfg) {
    record_error(e.id, e.msg.c_str());
    if (cfg == nullptr || cfg->parentid == mpStandalone) {
        fprintf(stdout, "\nMCX ERROR(%d):%s\n", e.id, e.msg.c_str());
        fflush(stdout);
        exit(e.id);
    }
}

GlibcRand::GlibcRand(unsigned int seed) {
    if (seed == 0) seed = 1;
    std::vector<unsigned int> st(344);
    int w = (int)seed;
    st[0] = (unsigned int)w;
    for (int i = 1; i < 31; i++) {
        /* 16807 * w mod (2^31 - 1), Schrage's decomposition */
        int hi = w / 127773, lo = w % 127773;
        w = 16807 * lo - 2836 * hi;
        if (w < 0) w += 2147483647;
        st[i] = (unsigned int)w;
    }
    for (int i = 31; i < 34; i++) st[i] = st[i - 31];
    for (int i = 34; i < 344; i++) st[i] = st[i - 31] + st[i - 3];
    for (int i = 0; i < 31; i++) ring_[i] = st[313 + i];
    pos_ = 0;
}

int GlibcRand::next() {
    unsigned int v = ring_[pos_] + ring_[(pos_ + 28) % 31];
    ring_[pos_] = v;
    pos_ = (pos_ + 1) % 31;
    return (int)(v >> 1);
}

void GlibcRand::state(uint32_t x[31]) const {
    for (int j = 0; j < 31; j++) x[j] = ring_[(pos_ + j) % 31];
}

void GlibcRand::set_state(const uint32_t x[31]) {
    for (int j = 0; j < 31; j++) ring_[j] = x[j];
    pos_ = 0;
}

namespace {
/* One output of the additive generator maps the window x of the 31 most
 * recent values (oldest first) to x' = A x over Z/2^32: x'[j] = x[j + 1]
 * for j < 30 and x'[30] = x[0] + x[28]. */
struct Jump {
    uint32_t m[31][31];
};

Jump jump_mul(const Jump& a, const Jump& b) {
    Jump c;
    memset(&c, 0, sizeof(c));
    for (int i = 0; i < 31; i++)
        for (int k = 0; k < 31; k++) {
            const uint32_t v = a.m[i][k];
            if (!v) continue;
            for (int j = 0; j < 31; j++) c.m[i][j] += v * b.m[k][j];
        }
    return c;
}

/* A^k: the window k outputs further on */
Jump jump_pow(uint64_t k) {
    Jump r, a;
    memset(&r, 0, sizeof(r));
    memset(&a, 0, sizeof(a));
    for (int i = 0; i < 31; i++) r.m[i][i] = 1;
    for (int j = 0; j < 30; j++) a.m[j][j + 1] = 1;
    a.m[30][0] = a.m[30][28] = 1;
    while (k) {
        if (k & 1) r = jump_mul(r, a);
        a = jump_mul(a, a);
        k >>= 1;
    }
    return r;
}

void jump_apply(const Jump& a, const uint32_t* x, uint32_t* y) {
    for (int i = 0; i < 31; i++) {
        uint32_t v = 0;
        for (int k = 0; k < 31; k++) v += a.m[i][k] * x[k];
        y[i] = v;
    }
}

/* the generator's values in blocks: v[i] = v[i - 31] + v[i - 3] over a
 * buffer whose first 31 entries are the window; rand() is v >> 1 */
void seed_words_seq(GlibcRand& r, uint32_t* out, size_t n) {
    const size_t kBlock = 1024; /* seed words per block */
    uint32_t v[31 + 3 * kBlock];
    r.state(v);
    for (size_t i0 = 0; i0 < n; i0 += kBlock) {
        const size_t m = std::min(kBlock, n - i0);
        for (size_t i = 31; i < 31 + 3 * m; i++) v[i] = v[i - 31] + v[i - 3];
        for (size_t i = 0; i < m; i++) {
            const uint32_t a = v[31 + 3 * i] >> 1, b = v[32 + 3 * i] >> 1, c = v[33 + 3 * i] >> 1;
            out[i0 + i] = (a << 16) | (b << 1) | (c >> 14);
        }
        memmove(v, v + 3 * m, 31 * sizeof(uint32_t));
    }
    r.set_state(v);
}
}  // namespace

void GlibcRand::seed_words(uint32_t* out, size_t n) {
    const size_t kChunks = 256;
    if (n < (size_t)1 << 18) {
        seed_words_seq(*this, out, n);
        return;
    }
    const size_t w = (n + kChunks - 1) / kChunks;
    const Jump step = jump_pow(3 * (uint64_t)w);
    std::vector<uint32_t> starts(31 * (kChunks + 1));
    state(starts.data());
    for (size_t c = 0; c < kChunks; c++) jump_apply(step, &starts[31 * c], &starts[31 * (c + 1)]);
#pragma omp parallel for schedule(dynamic)
    for (long long c = 0; c < (long long)kChunks; c++) {
        const size_t b = (size_t)c * w, e = std::min(n, b + w);
        if (b >= e) continue;
        GlibcRand g(1u);
        g.set_state(&starts[31 * c]);
        seed_words_seq(g, out + b, e - b);
    }
    uint32_t fin[31];
    jump_apply(jump_pow(3 * (uint64_t)n), &starts[0], fin);
    set_state(fin);
}

unsigned int det_partialdata(const Config* cfg) {
    const unsigned int f = cfg->savedetflag;
    return (cfg->medianum - 1) * (MCX_SAVE_NSCAT(f) + MCX_SAVE_PPATH(f) + MCX_SAVE_MOM(f));
}

unsigned int det_reclen(const Config* cfg) {
    const unsigned int f = cfg->savedetflag;
    return det_partialdata(cfg) + MCX_SAVE_DETID(f) + 3 * (MCX_SAVE_PEXIT(f) + MCX_SAVE_VEXIT(f)) + MCX_SAVE_W0(f) +
           4 * MCX_SAVE_IQUV(f);
}

/* position of c in a NUL-terminated table, or -1 */
static int lookup_index(char c, const char* table) {
    for (int i = 0; table[i]; i++)
        if (tolower((unsigned char)c) == table[i]) return i;
    return -1;
}

static void maskdet_impl(Config* cfg) {
    const float corners[8][3] = {{0, 0, 0}, {1, 0, 0}, {0, 1, 0}, {0, 0, 1}, {1, 1, 0}, {1, 0, 1}, {0, 1, 1}, {1, 1, 1}};
    const unsigned int nx = cfg->dim.x, ny = cfg->dim.y, nz = cfg->dim.z;
    const size_t dx = nx + 2, dy = ny + 2, dz = nz + 2;
    const int isonecube = (nx == 1) && (ny == 1) && (nz == 1);
    const bool svmc = cfg->mediabyte == MCX_MEDIA_2LABEL_SPLIT;
    /* a zero-padded copy, so the 26-neighbour test needs no bounds checks */
    std::vector<unsigned int> pad(dx * dy * dz, 0);
    for (size_t z = 1; z <= nz; z++)
        for (size_t y = 1; y <= ny; y++)
            memcpy(&pad[z * dy * dx + y * dx + 1], cfg->vol + (z - 1) * ny * nx + (y - 1) * nx, nx * sizeof(unsigned int));
    for (unsigned int d = 0; d < cfg->detnum; d++) {
        const mcx_float4 dp = cfg->detpos[d];
        unsigned int count = 0;
        const float radius = fabsf(dp.w);
        const float d2max = (radius + 1.7321f) * (radius + 1.7321f);
        for (float z = -radius - 1.f; z <= radius + 1.f; z += 0.5f) {
            const float iz = z + dp.z;
            for (float y = -radius - 1.f; y <= radius + 1.f; y += 0.5f) {
                const float iy = y + dp.y;
                for (float x = -radius - 1.f; x <= radius + 1.f; x += 0.5f) {
                    const float ix = x + dp.x;
                    if (iz < 0 || ix < 0 || iy < 0 || ix >= nx || iy >= ny || iz >= nz ||
                        x * x + y * y + z * z > (radius + 1.f) * (radius + 1.f))
                        continue;
                    float mind2 = MCX_VERY_BIG;
                    for (int c = 0; c < 8; c++) {
                        const float rx = (int)ix - dp.x + corners[c][0];
                        const float ry = (int)iy - dp.y + corners[c][1];
                        const float rz = (int)iz - dp.z + corners[c][2];
                        const float d2 = rx * rx + ry * ry + rz * rz;
                        if (d2 > d2max) {
                            mind2 = MCX_VERY_BIG;
                            break;
                        }
                        if (d2 < mind2) mind2 = d2;
                    }
                    const float lim = radius + 0.5f * (1.f + isonecube);
                    if (mind2 == MCX_VERY_BIG || mind2 >= lim * lim) continue;
                    const size_t i = (size_t)(int)(iz + 1.f) * dy * dx + (size_t)(int)(iy + 1.f) * dx + (size_t)(int)(ix + 1.f);
                    /* split voxels (src/mcx_utils.c:4287-4310): a voxel whose
                     * lower label is the background is on the surface; one
                     * with a tissue lower label is when a neighbour's lower
                     * label is the background */
                    const unsigned int lower = (pad[i] & 0xFF000000u) >> 24, upper = (pad[i] & 0x00FF0000u) >> 16;
                    if (svmc ? !(lower || upper) : !pad[i]) continue;
                    bool surrounded = !(svmc && !lower && upper);
                    for (int oz = -1; oz <= 1 && surrounded; oz++)
                        for (int oy = -1; oy <= 1 && surrounded; oy++)
                            for (int ox = -1; ox <= 1; ox++) {
                                if (!ox && !oy && !oz) continue;
                                const size_t j = (size_t)((long)i + (long)oz * (long)(dy * dx) + (long)oy * (long)dx + ox);
                                if (svmc ? !((pad[j] & 0xFF000000u) >> 24) : !pad[j]) {
                                    surrounded = false;
                                    break;
                                }
                            }
                    if (!surrounded) {
                        cfg->vol[(size_t)(int)iz * ny * nx + (size_t)(int)iy * nx + (int)ix] |= MCX_DET_MASK;
                        count++;
                    }
                }
            }
        }
        if (cfg->issavedet && count == 0)
            fprintf(stderr, "WARNING: detector %u is not located on an interface, please check coordinates.\n", d + 1);
    }
}

/* src/mcx_utils.c:5490-5546 (mcx_float2half2): the reference's float ->
 * half encoder, reproduced bit for bit because it defines the words that
 * the kernel decodes.  Its rules: the mantissa is truncated to 10 bits; an
 * exponent at or below 2^-15 gives a zero half exponent, except that when
 * the top mantissa bits are also small (below 16) the value takes a
 * rounded denormal path.  Because the encoder writes the first half into
 * the union word it is still reading, the first value's three lowest
 * mantissa bits are taken from the half just written: exponent bit 3,
 * exponent bit 4 and the sign, not float bits 13-15.  A packed word of 0
 * becomes 0x00010000 (mus = the smallest half denormal) and a packed word
 * of 0x80000000 becomes 0.  For exponents below 2^-45 the denormal path of
 * the reference shifts by 32 or more (undefined in C); here such values
 * encode to a signed zero. */
static unsigned int half_exp(unsigned int b) {
    const unsigned int e = (b >> 23) & 0xFFu;
    return e > 0x70u ? ((e - 0x70u) & 0x1Fu) : 0u;
}

static unsigned int half_denorm(unsigned int b) {
    const unsigned int e = (b >> 23) & 0xFFu;
    const unsigned int m = ((b >> 12) & 0x7FFu) | 0x800u;
    const unsigned int s1 = 114u - e, s2 = 113u - e;
    const unsigned int q = (s1 < 32u ? (m >> s1) : 0u) + (s2 < 32u ? ((m >> s2) & 1u) : 0u);
    return ((b >> 16) & 0x8000u) | (q & 0xFFFFu);
}

unsigned int encode_half2(float a, float b) {
    unsigned int ba, bb;
    memcpy(&ba, &a, 4);
    memcpy(&bb, &b, 4);
    unsigned int h0, h1;
    const unsigned int t0 = half_exp(ba), t1 = half_exp(bb);
    if (((ba >> 13) & 0x3FFu) < 0x10u && t0 == 0u) {
        h0 = half_denorm(ba);
    } else {
        const unsigned int top = ((((ba >> 31) << 5) | t0) << 10) & 0xFFFFu;
        h0 = top | ((ba >> 13) & 0x3F8u) | ((top >> 13) & 7u);
    }
    if (((bb >> 13) & 0x3FFu) < 0x10u && t1 == 0u) {
        h1 = half_denorm(bb);
    } else {
        h1 = (((((bb >> 31) << 5) | t1) << 10) & 0xFFFFu) | ((bb >> 13) & 0x3FFu);
    }
    unsigned int w = (h0 & 0xFFFFu) | (h1 << 16);
    if (w == 0u) w = 0x00010000u;
    if (w == 0x80000000u) w = 0u;
    return w;
}

/* src/mcx_utils.c:3551-3569: a float mua scaled to grid units; 0 becomes
 * FLT_EPSILON (a 0 word is a void voxel), NaN becomes a void voxel, and
 * (volume files only) so does -0 */
unsigned int encode_mua_float(float mua, float unitinmm, bool signed_zero_is_void) {
    float f = mua * unitinmm;
    unsigned int w;
    memcpy(&w, &f, 4);
    if (w == 0u) {
        f = MCX_EPS;
        memcpy(&w, &f, 4);
    }
    if (mua != mua || (signed_zero_is_void && w == 0x80000000u)) w = 0u;
    return w;
}

/* src/mcx_utils.c:1714-1738: split voxels arrive as 8 bytes per voxel
 * {lower label, upper label, point x, y, z, normal x, y, z}; the kernel
 * reads them as two planes of words, the first holding
 * {point y, point x, upper, lower} from the low byte up and the second
 * (nvox words on) {normal z, normal y, normal x, point z} */
static void svmc_planes(unsigned int* vol, size_t n) {
    std::vector<unsigned int> out(2 * n);
    const unsigned char* b = (const unsigned char*)vol;
    for (size_t i = 0; i < n; i++) {
        const unsigned char* v = b + 8 * i;
        out[i] = (unsigned int)v[3] | ((unsigned int)v[2] << 8) | ((unsigned int)v[1] << 16) | ((unsigned int)v[0] << 24);
        out[i + n] = (unsigned int)v[7] | ((unsigned int)v[6] << 8) | ((unsigned int)v[5] << 16) | ((unsigned int)v[4] << 24);
    }
    memcpy(vol, out.data(), 8 * n);
}

/* src/mcx_utils.c:1622-1623 (mcx_convertrow2col64): 8-byte voxels */
static void row2col64(unsigned int* vol, const mcx_uint3* dim) {
    const size_t n = (size_t)dim->x * dim->y * dim->z;
    std::vector<unsigned long long> tmp(n);
    memcpy(tmp.data(), vol, 8 * n);
    unsigned long long* v = (unsigned long long*)vol;
    for (size_t x = 0; x < dim->x; x++)
        for (size_t y = 0; y < dim->y; y++)
            for (size_t z = 0; z < dim->z; z++)
                v[z * dim->y * dim->x + y * dim->x + x] = tmp[x * dim->y * dim->z + y * dim->z + z];
}

static void row2col(unsigned int* vol, const mcx_uint3* dim) {
    const size_t n = (size_t)dim->x * dim->y * dim->z;
    std::vector<unsigned int> tmp(vol, vol + n);
    for (size_t x = 0; x < dim->x; x++)
        for (size_t y = 0; y < dim->y; y++)
            for (size_t z = 0; z < dim->z; z++)
                vol[z * dim->y * dim->x + y * dim->x + x] = tmp[x * dim->y * dim->z + y * dim->z + z];
}

static void preprocess_impl(Config* cfg) {
    int isbcdet = 0;
    double tmp = sqrt((double)cfg->srcdir.x * cfg->srcdir.x + (double)cfg->srcdir.y * cfg->srcdir.y +
                      (double)cfg->srcdir.z * cfg->srcdir.z);
    if (tmp < MCX_EPS) MCX_THROW(-4, "source initial direction vector can not have a length of 0");
    tmp = 1.f / tmp;
    cfg->srcdir.x *= tmp;
    cfg->srcdir.y *= tmp;
    cfg->srcdir.z *= tmp;
    if (cfg->debuglevel & MCX_DEBUG_MOVE_ONLY) {
        cfg->issave2pt = 0;
        cfg->issavedet = 0;
    }
    if ((cfg->outputtype == otJacobian || cfg->outputtype == otWP || cfg->outputtype == otDCS ||
         cfg->outputtype == otRF || cfg->outputtype == otRFmus || cfg->outputtype == otWLTOF ||
         cfg->outputtype == otWPTOF) && cfg->seed != MCX_SEED_FROM_FILE)
        MCX_THROW(-6, "Jacobian output is only valid in the reply mode. Please define cfg.seed");
    cfg->istrajstokes = 0;
    static const char bccodes[] = {'_', 'r', 'a', 'm', 'c', '\0'};
    static const char bcdet[] = {'0', '1', '\0'};
    for (int i = 0; i < 6; i++) {
        if (cfg->bc[i] && cfg->bc[i] > 4) {
            int k = lookup_index(cfg->bc[i], bccodes);
            if (k < 0) MCX_THROW(-4, "unknown boundary condition specifier");
            cfg->bc[i] = (char)k;
        }
    }
    for (int i = 6; i < 12; i++) {
        if (cfg->bc[i] && cfg->bc[i] > 1) {
            int k = lookup_index(cfg->bc[i], bcdet);
            if (k < 0) MCX_THROW(-4, "unknown boundary detection flags");
            cfg->bc[i] = (char)k;
        }
        if (cfg->bc[i]) isbcdet = 1;
    }
    if (cfg->vol && cfg->polprop) MCX_THROW(-1, "polarized photon simulation is not supported by this engine");
    if (cfg->medianum == 0) MCX_THROW(-4, "you must define the 'prop' field in the input structure");
    if (cfg->dim.x == 0 || cfg->dim.y == 0 || cfg->dim.z == 0)
        MCX_THROW(-4, "the 'vol' field in the input structure can not be empty");
    if ((cfg->srctype == MCX_SRC_PATTERN || cfg->srctype == MCX_SRC_PATTERN3D) && cfg->srcpattern == NULL)
        MCX_THROW(-4, "the 'srcpattern' field can not be empty when your 'srctype' is 'pattern'");
    if (cfg->steps.x != 1.f && cfg->unitinmm == 1.f) cfg->unitinmm = cfg->steps.x;
    if (cfg->unitinmm != 1.f) {
        cfg->steps.x = cfg->steps.y = cfg->steps.z = cfg->unitinmm;
        for (unsigned int i = 1; i < cfg->medianum; i++) {
            cfg->prop[i].mus *= cfg->unitinmm;
            cfg->prop[i].mua *= cfg->unitinmm;
        }
    }
    if (cfg->isrowmajor) {
        /* src/mcx_utils.c:1620-1630: ASGN_F2H keeps two planes of words,
         * split voxels are 8-byte units */
        if (cfg->mediabyte == MCX_MEDIA_2LABEL_SPLIT) {
            row2col64(cfg->vol, &cfg->dim);
        } else {
            row2col(cfg->vol, &cfg->dim);
            if (cfg->mediabyte == MCX_MEDIA_ASGN_F2H) row2col(cfg->vol + (size_t)cfg->dim.x * cfg->dim.y * cfg->dim.z, &cfg->dim);
        }
        cfg->isrowmajor = 0;
    }
    if (cfg->issavedet && cfg->detnum == 0 && isbcdet == 0) cfg->issavedet = 0;
    if (cfg->issavedet == 0) {
        cfg->issaveexit = 0;
        cfg->ismomentum = 0;
        if (cfg->seed != MCX_SEED_FROM_FILE) cfg->savedetflag = 0;
    }
    if (cfg->respin == 0) MCX_THROW(-1, "respin number can not be 0, check your -r/--repeat input or cfg.respin value");
    if (cfg->replaydet > (int)cfg->detnum) MCX_THROW(-4, "replay detector ID exceeds the maximum detector number");
    if (cfg->replaydet == -1 && cfg->detnum == 1) cfg->replaydet = 1;
    for (unsigned int i = 0; i < cfg->medianum; i++) {
        if (cfg->prop[i].mus == 0.f) {
            cfg->prop[i].mus = MCX_EPS;
            cfg->prop[i].g = 1.f;
        }
    }
    if (cfg->srcpos.w == 0.f) cfg->srcpos.w = 1.f;
    if (cfg->extrasrclen) MCX_THROW(-4, "multiple sources (srcdata) are not supported by this engine");
    if (cfg->vol) {
        const size_t dimxyz = (size_t)cfg->dim.x * cfg->dim.y * cfg->dim.z;
        unsigned int maxlabel = 0;
        if (cfg->mediabyte <= 4) {
            for (size_t i = 0; i < dimxyz; i++) maxlabel = std::max(maxlabel, cfg->vol[i] & MCX_MED_MASK);
        } else if (cfg->mediabyte == MCX_MEDIA_2LABEL_MIX) {
            /* the two byte labels of each word index the property table */
            for (size_t i = 0; i < dimxyz; i++)
                maxlabel = std::max(maxlabel, std::max(cfg->vol[i] & 0xFFu, (cfg->vol[i] >> 8) & 0xFFu));
        } else if (cfg->mediabyte == MCX_MEDIA_LABEL_HALF) {
            for (size_t i = 0; i < dimxyz; i++) maxlabel = std::max(maxlabel, cfg->vol[i] & 0x3FFFu);
        } else if (cfg->mediabyte == MCX_MEDIA_2LABEL_SPLIT) {
            svmc_planes(cfg->vol, dimxyz);
            for (size_t i = 0; i < dimxyz; i++) {
                if (cfg->vol[i] >> 31) MCX_THROW(-4, "split-voxel lower labels must be below 128");
                maxlabel = std::max(maxlabel, std::max(cfg->vol[i] >> 24, (cfg->vol[i] >> 16) & 0xFFu));
            }
        }
        if (cfg->medianum <= maxlabel) MCX_THROW(-4, "input media optical properties are less than the labels in the volume");
    }
    /* point sources: precompute the launch voxel index and its label
     * into srcparam2.z/.w (bit patterns) */
    if (cfg->srctype <= MCX_SRC_CONE || cfg->srctype == MCX_SRC_ARCSINE || cfg->srctype == MCX_SRC_ZGAUSSIAN) {
        unsigned int i1 = 0, m1 = 0;
        if (!(cfg->srcpos.x < 0.f || cfg->srcpos.y < 0.f || cfg->srcpos.z < 0.f || cfg->srcpos.x >= cfg->dim.x ||
              cfg->srcpos.y >= cfg->dim.y || cfg->srcpos.z >= cfg->dim.z)) {
            i1 = (unsigned int)((int)floorf(cfg->srcpos.z) * (cfg->dim.y * cfg->dim.x) +
                                (int)floorf(cfg->srcpos.y) * cfg->dim.x + (int)floorf(cfg->srcpos.x));
            m1 = cfg->vol[i1] & MCX_MED_MASK;
        }
        memcpy(&cfg->srcparam2.z, &i1, 4);
        memcpy(&cfg->srcparam2.w, &m1, 4);
    }
    if (cfg->issavedet) maskdet_impl(cfg);
    for (int i = 0; i < MCX_MAX_DEVICE; i++)
        if (cfg->deviceid[i] == '0') cfg->deviceid[i] = '\0';
    if ((cfg->mediabyte == MCX_MEDIA_AS_F2H || cfg->mediabyte == MCX_MEDIA_MUA_FLOAT || cfg->mediabyte == MCX_MEDIA_AS_HALF) &&
        cfg->medianum < 2)
        MCX_THROW(-4, "the 'prop' field must contain at least 2 rows for the requested media format");
    if ((cfg->mediabyte == MCX_MEDIA_ASGN_BYTE || cfg->mediabyte == MCX_MEDIA_AS_SHORT) && cfg->medianum < 3)
        MCX_THROW(-4, "the 'prop' field must contain at least 3 rows for the requested media format");
    if (cfg->ismomentum) cfg->savedetflag |= (1u << 3);
    if (cfg->issaveexit) cfg->savedetflag |= (1u << 4) | (1u << 5);
    cfg->savedetflag &= ~(1u << 7);
    if (cfg->issavedet && cfg->savedetflag == 0) cfg->savedetflag = 0x5;
    /* src/mcx_utils.c:1801-1806: property-word media have no medium index
     * for the per-medium record columns (scattering counts, partial paths,
     * momentum transfer).  The reference drops them from 100 on; here the
     * formats below 100 drop them too, since their words carry no index
     * that fits the columns either. */
    if (cfg->mediabyte > 4 && cfg->mediabyte != MCX_MEDIA_2LABEL_SPLIT && cfg->savedetflag)
        cfg->savedetflag &= ~((1u << 1) | (1u << 2) | (1u << 3) | (1u << 7));
    if (cfg->issaveref > 1) {
        if (cfg->issavedet == 0) MCX_THROW(-4, "you must have at least two outputs if issaveref is greater than 1");
        cfg->savedetflag = 0x5;
    }
}

}  // namespace mcx

using namespace mcx;

extern "C" {

void mcx_initcfg(Config* cfg) {
    memset(cfg, 0, sizeof(Config));
    cfg->medianum = 0;
    cfg->mediabyte = 1;
    cfg->steps.x = cfg->steps.y = cfg->steps.z = 1.f;
    cfg->nblocksize = 64;
    cfg->nthread = (1 << 14);
    cfg->isreflect = 1;
    cfg->isref3 = 1;
    cfg->isnormalized = 1;
    cfg->issavedet = 1;
    cfg->respin = 1;
    cfg->issave2pt = 1;
    cfg->flog = stdout;
    cfg->sradius = -2.f;
    cfg->tstep = MCX_VERY_BIG;
    cfg->unitinmm = 1.f;
    cfg->maxdetphoton = 1000000;
    cfg->maxjumpdebug = 10000000;
    cfg->autopilot = 1;
    cfg->seed = 0x623F9A9E;
    memcpy(cfg->his.magic, "MCXH", 4);
    cfg->his.version = 1;
    cfg->his.unitinmm = 1.f;
    cfg->his.normalizer = 1.f;
    cfg->his.respin = 1;
    cfg->his.srcnum = 1;
    cfg->savedetflag = 0x5;
    cfg->his.savedetflag = cfg->savedetflag;
    cfg->maxvoidstep = 1000;
    cfg->voidtime = 1;
    cfg->srcnum = 1;
    cfg->outputtype = otFlux;
    cfg->outputformat = ofJNifti;
    cfg->srcpos.w = 1.f;
    cfg->gscatter = 1000000000u;
    cfg->his.totalsource = 1;
    cfg->deviceid[0] = '1';
    cfg->parentid = mpStandalone;
}

void mcx_clearcfg(Config* cfg) {
    if (cfg->medianum) free(cfg->prop);
    if (cfg->polmedianum) free(cfg->polprop);
    if (cfg->smatrix) free(cfg->smatrix);
    if (cfg->detnum) free(cfg->detpos);
    if (cfg->dim.x && cfg->dim.y && cfg->dim.z) free(cfg->vol);
    if (cfg->srcpattern) free(cfg->srcpattern);
    if (cfg->replay.weight) free(cfg->replay.weight);
    if (cfg->replay.seed) free(cfg->replay.seed);
    if (cfg->replay.tof) free(cfg->replay.tof);
    if (cfg->replay.detid) free(cfg->replay.detid);
    if (cfg->exportfield) free(cfg->exportfield);
    if (cfg->exportdetected) free(cfg->exportdetected);
    if (cfg->exportdebugdata) free(cfg->exportdebugdata);
    if (cfg->seeddata) free(cfg->seeddata);
    if (cfg->shapedata) free(cfg->shapedata);
    if (cfg->extrajson) free(cfg->extrajson);
    if (cfg->invcdf) free(cfg->invcdf);
    if (cfg->angleinvcdf) free(cfg->angleinvcdf);
    if (cfg->srcdata) free(cfg->srcdata);
    if (cfg->dx) free(cfg->dx);
    if (cfg->dy) free(cfg->dy);
    if (cfg->dz) free(cfg->dz);
    if (cfg->flog && cfg->flog != stdout && cfg->flog != stderr) fclose(cfg->flog);
    mcx_initcfg(cfg);
}

void mcx_cleargpuinfo(GPUInfo** gpuinfo) {
    if (*gpuinfo) {
        free(*gpuinfo);
        *gpuinfo = NULL;
    }
}

void mcx_error(const int id, const char* msg, const char* file, const int linenum) {
    if (g_depth > 0) throw mcx::Error(id, msg ? msg : "");
    record_error(id, msg);
    fprintf(stdout, "\nMCX ERROR(%d):%s in unit %s:%d\n", id, msg ? msg : "", file ? file : "", linenum);
    fflush(stdout);
    exit(id);
}

int mcx_last_error(char* msg, int msglen) {
    std::lock_guard<std::mutex> lk(g_errmu);
    if (msg && msglen > 0) {
        strncpy(msg, g_errmsg.c_str(), (size_t)msglen - 1);
        msg[msglen - 1] = '\0';
    }
    return g_errid;
}

void mcx_clear_error(void) {
    std::lock_guard<std::mutex> lk(g_errmu);
    g_errid = 0;
    g_errmsg.clear();
}

void mcx_flush(Config* cfg) {
    if (cfg && cfg->flog) fflush(cfg->flog);
}

int mcx_float2half2(float input[2]) { return (int)encode_half2(input[0], input[1]); }

void mcx_normalize(float field[], float scale, size_t fieldlen, int option, int pidx, int srcnum) {
#pragma omp parallel for schedule(static)
    for (long long i = 0; i < (long long)fieldlen; i++) {
        float& f = field[(size_t)i * srcnum + pidx];
        if (option == 2 && f < 0.f) continue;
        f *= scale;
    }
}

void mcx_maskdet(Config* cfg) {
    EntryGuard g;
    try {
        maskdet_impl(cfg);
    } catch (const Error& e) {
        handle_error(e, cfg);
    }
}

void mcx_preprocess(Config* cfg) {
    EntryGuard g;
    try {
        preprocess_impl(cfg);
    } catch (const Error& e) {
        handle_error(e, cfg);
    }
}

/* src/mcx_utils.c:3624-3699: select the detected photons to replay (all, or
 * those of detector replaydet), compacting their seeds, and derive each
 * one's detected weight exp(-sum mua_m L_m) and time of flight
 * sum n_m L_m / c0 from its partial paths; photons whose time of flight lies
 * outside [tstart, tend] are dropped.  detps holds one record of
 * dimdetps[0] floats per photon (dimdetps[1] photons). */
static void replay_init(Config* cfg, const float* detps, const int* dimdetps, int seedbyte) {
    if (cfg->seed == MCX_SEED_FROM_FILE && detps == NULL)
        MCX_THROW(-6, "you give cfg.seed for replay, but did not specify cfg.detphotons.\nPlease define it as the detphoton output from the baseline simulation\n");
    if (detps == NULL || cfg->seed != MCX_SEED_FROM_FILE) return;
    if (!dimdetps || cfg->nphoton != (size_t)dimdetps[1]) MCX_THROW(-6, "the column numbers of detphotons and seed do not match\n");
    if (seedbyte == 0 || !cfg->replay.seed) MCX_THROW(-6, "the seed input is empty");
    if (seedbyte != 16) MCX_THROW(-6, "the seed byte-length does not match this engine's RNG (16 bytes)");
    const int hasdetid = cfg->savedetflag & 1;
    const int offset = ((cfg->savedetflag >> 1) & 1) * (int)(cfg->medianum - 1);
    if ((!hasdetid && cfg->detnum > 1) || !((cfg->savedetflag >> 2) & 1))
        MCX_THROW(-6, "please rerun the baseline simulation and save detector ID (D) and partial-path (P) using cfg.savedetflag='dp' ");
    const int n = dimdetps[1], rl = dimdetps[0];
    if (rl < hasdetid + offset + (int)cfg->medianum - 1) MCX_THROW(-6, "detphotons records are shorter than savedetflag implies");
    if (cfg->replay.weight) free(cfg->replay.weight);
    if (cfg->replay.tof) free(cfg->replay.tof);
    if (cfg->replay.detid) free(cfg->replay.detid);
    cfg->replay.weight = (float*)malloc(std::max(n, 1) * sizeof(float));
    cfg->replay.tof = (float*)calloc(std::max(n, 1), sizeof(float));
    cfg->replay.detid = (int*)calloc(std::max(n, 1), sizeof(int));
    if (!cfg->replay.weight || !cfg->replay.tof || !cfg->replay.detid) MCX_THROW(-1, "out of memory for replay data");
    char* seeds = (char*)cfg->replay.seed;
    size_t kept = 0;
    for (int i = 0; i < n; i++) {
        const float* rec = detps + (size_t)i * rl;
        const int det = hasdetid ? (int)rec[0] : 1;
        if (cfg->replaydet > 0 && cfg->replaydet != (det & 0xFFFF)) continue;
        if (cfg->srcid > 0 && !((det & 0xFFFF0000) && cfg->srcid == ((det & 0xFFFF0000) >> 16))) continue;
        float w = 1.f, tof = 0.f;
        for (unsigned int m = 0; m + 1 < cfg->medianum; m++) {
            float plen = rec[hasdetid + offset + m];
            w *= expf(-cfg->prop[m + 1].mua * plen);
            plen *= cfg->unitinmm;
            tof += plen * MCX_R_C0 * cfg->prop[m + 1].n;
        }
        if (tof < cfg->tstart || tof > cfg->tend) continue;
        if ((size_t)i != kept) memmove(seeds + kept * seedbyte, seeds + (size_t)i * seedbyte, seedbyte);
        cfg->replay.weight[kept] = w;
        cfg->replay.tof[kept] = tof;
        cfg->replay.detid[kept] = det;
        kept++;
    }
    cfg->nphoton = kept;
    cfg->his.seedbyte = seedbyte;
}

void mcx_validatecfg(Config* cfg, float* detps, int dimdetps[2], int seedbyte) {
    EntryGuard g;
    try {
        unsigned int hostdetreclen = det_reclen(cfg);
        if (!cfg->issave2pt && cfg->issaveref) cfg->issaveref = 0;
        if (!cfg->issrcfrom0) {
            cfg->srcpos.x--;
            cfg->srcpos.y--;
            cfg->srcpos.z--;
        }
        if (cfg->tstep == 0.f) cfg->tstep = cfg->tend;
        if (cfg->tstart >= cfg->tend || cfg->tstep == 0.f) MCX_THROW(-6, "incorrect time gate settings");
        if (fabs(cfg->srcdir.x * cfg->srcdir.x + cfg->srcdir.y * cfg->srcdir.y + cfg->srcdir.z * cfg->srcdir.z - 1.f) > 1e-5)
            MCX_THROW(-6, "field 'srcdir' must be a unitary vector");
        if (cfg->steps.x == 0.f || cfg->steps.y == 0.f || cfg->steps.z == 0.f)
            MCX_THROW(-6, "field 'steps' can not have zero elements");
        unsigned int gates = (unsigned int)((cfg->tend - cfg->tstart) / cfg->tstep + 0.5);
        if (cfg->maxgate > gates) cfg->maxgate = gates;
        if (cfg->seed < 0 && cfg->seed != MCX_SEED_FROM_FILE) cfg->seed = (int)time(NULL);
        for (unsigned int i = 0; i < cfg->detnum; i++) {
            if (!cfg->issrcfrom0) {
                cfg->detpos[i].x--;
                cfg->detpos[i].y--;
                cfg->detpos[i].z--;
            }
        }
        if (cfg->shapedata) apply_shapes(cfg, cfg->shapedata);
        preprocess_impl(cfg);
        cfg->his.maxmedia = cfg->medianum - 1;
        cfg->his.detnum = cfg->detnum;
        cfg->his.srcnum = cfg->srcnum;
        cfg->his.colcount = det_reclen(cfg);
        (void)hostdetreclen;
        cfg->his.savedetflag = cfg->savedetflag;
        if (cfg->issaveseed) cfg->his.seedbyte = 16;
        replay_init(cfg, detps, dimdetps, seedbyte);
    } catch (const Error& e) {
        handle_error(e, cfg);
    }
}

}  // extern "C"

/* Byte offsets of the reference's Config/GPUInfo/History (src/mcx_utils.h:
 * 115-285) under the x86-64 C ABI with 16-byte aligned float4, derived field
 * by field from the header; a mismatch here breaks drop-in callers. */
static_assert(sizeof(History) == 64, "History layout");
static_assert(sizeof(GPUInfo) == 328, "GPUInfo layout");
static_assert(offsetof(GPUInfo, globalmem) == 272 && offsetof(GPUInfo, maxgate) == 320, "GPUInfo layout");
static_assert(offsetof(Config, srcpos) == 32 && offsetof(Config, tstart) == 80 && offsetof(Config, steps) == 92, "Config layout");
static_assert(offsetof(Config, dim) == 104 && offsetof(Config, medianum) == 140 && offsetof(Config, prop) == 168, "Config layout");
static_assert(offsetof(Config, maxgate) == 200 && offsetof(Config, vol) == 216 && offsetof(Config, session) == 224, "Config layout");
static_assert(offsetof(Config, isrowmajor) == 480 && offsetof(Config, zipid) == 500 && offsetof(Config, srctype) == 504, "Config layout");
static_assert(offsetof(Config, minenergy) == 508 && offsetof(Config, flog) == 528 && offsetof(Config, his) == 536, "Config layout");
static_assert(offsetof(Config, exportfield) == 600 && offsetof(Config, detectedcount) == 616 && offsetof(Config, rootpath) == 624, "Config layout");
static_assert(offsetof(Config, maxvoidstep) == 1664 && offsetof(Config, srcparam1) == 1680 && offsetof(Config, srcnum) == 1712, "Config layout");
static_assert(offsetof(Config, replay) == 1728 && offsetof(Config, seeddata) == 1760 && offsetof(Config, seedfile) == 1772, "Config layout");
static_assert(offsetof(Config, debuglevel) == 3820 && offsetof(Config, deviceid) == 3828 && offsetof(Config, workload) == 4084, "Config layout");
static_assert(offsetof(Config, parentid) == 5108 && offsetof(Config, energytot) == 5152 && offsetof(Config, normalizer) == 5176, "Config layout");
static_assert(offsetof(Config, mediabyte) == 5200 && offsetof(Config, bc) == 5232 && offsetof(Config, nphase) == 5248, "Config layout");
static_assert(offsetof(Config, srcid) == 5280 && offsetof(Config, srcdata) == 5288 && sizeof(Config) == 5296, "Config layout");

extern "C" int mcx_amd_abi_layout(size_t* out, int n) {
    const size_t v[] = {sizeof(Config), sizeof(GPUInfo), sizeof(History), offsetof(Config, srcpos), offsetof(Config, session),
                        offsetof(Config, flog), offsetof(Config, his), offsetof(Config, exportfield), offsetof(Config, srcparam1),
                        offsetof(Config, deviceid), offsetof(Config, energytot), offsetof(Config, bc), offsetof(Config, srcdata),
                        sizeof(McxBatchOptions), offsetof(McxBatchOptions, photons)};
    int k = n < (int)(sizeof(v) / sizeof(v[0])) ? n : (int)(sizeof(v) / sizeof(v[0]));
    for (int i = 0; i < k; i++) out[i] = v[i];
    return (int)(sizeof(v) / sizeof(v[0]));
}
