/*
 * mcx_io.cpp -- JSON input surface and output writers.
 *
 *   mcx_readconfig / JSON loader   src/mcx_utils.c:1418-1489, :2257-3210
 *   Shapes rasterizer              src/mcx_shapes.c:163-863
 *   mcx_run_from_json              src/mcx_utils.c:5578-5603
 *   mcx_savedata (.jnii, .mc2)     src/mcx_utils.c:916-970, :791-905, :4538-4614
 *   mcx_savedetphoton (.mch, .jdat) src/mcx_utils.c:982-1201
 *
 * .jnii stores the column-major float32 volume zlib-compressed and base64
 * encoded under NIFTIData._ArrayZipData_ with _ArrayOrder_ "c"; .mch is the
 * 64-byte History header followed by float32 records.
 */
#include <ctype.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <zlib.h>

#include <fstream>
#include <sstream>
#include <vector>

#include "mcx_internal.h"
#include "mcx_json.h"

namespace mcx {

static const char* kSrcTypes[] = {"pencil", "isotropic", "cone", "gaussian", "planar", "pattern", "fourier", "arcsine", "disk",
                                  "fourierx", "fourierx2d", "zgaussian", "line", "slit", "pencilarray", "pattern3d",
                                  "hyperboloid", "ring", nullptr};
static const char* kOutFormats[] = {"mc2", "nii", "hdr", "ubj", "tx3", "jnii", "bnii", nullptr};
static const char kOutTypes[] = {'x', 'f', 'e', 'j', 'p', 'm', 'r', 'l', 's', 't', 'b', '\0'};
static const char kSaveFlags[] = {'D', 'S', 'P', 'M', 'X', 'V', 'W', 'I', '\0'};
static const char kDebugFlags[] = {'R', 'M', 'P', 'T', '\0'};
static const char* kMediaFormats[] = {"byte", "short", "integer", "asgn_float", "svmc", "mixlabel", "labelplus",
                                      "muamus_float", "mua_float", "muamus_half", "asgn_byte", "muamus_short", nullptr};
static const int kMediaFormatIds[] = {1, 2, 4, 96, 97, 98, 99, 100, 101, 102, 103, 104};

static int key_lookup(const std::string& k, const char** table) {
    for (int i = 0; table[i]; i++)
        if (strcasecmp(k.c_str(), table[i]) == 0) return i;
    return -1;
}

/* letters of `s` -> bit mask by their position in `table` (case-insensitive) */
static unsigned int parse_flags(const std::string& s, const char* table) {
    unsigned int m = 0;
    for (char c : s)
        for (int i = 0; table[i]; i++)
            if (toupper((unsigned char)c) == table[i]) m |= (1u << i);
    return m;
}

static double num_at(const Json& a, size_t i) {
    if (!a.is_arr() || i >= a.arr.size()) throw Error(-1, "JSON: numeric array too short");
    return a.arr[i].as_num();
}

/* ---- Shapes (src/mcx_shapes.c) ---------------------------------------- */

struct Grid {
    unsigned int** vol;
    mcx_uint3* dim;
    float ox, oy, oz;
};

static inline size_t vidx(const Grid& g, int i, int j, int k) {
    return (size_t)k * g.dim->x * g.dim->y + (size_t)j * g.dim->x + i;
}

static int tag_of(const Json& o) {
    const Json* t = o.get("Tag");
    return t ? (int)t->as_num() : 0;
}

static void shape_grid(const Json& o, Grid& g) {
    const Json* sz = o.get("Size");
    if (!sz || sz->size() != 3) throw Error(-1, "A Grid command misses Size field");
    g.dim->x = (unsigned)num_at(*sz, 0);
    g.dim->y = (unsigned)num_at(*sz, 1);
    g.dim->z = (unsigned)num_at(*sz, 2);
    size_t n = (size_t)g.dim->x * g.dim->y * g.dim->z;
    free(*g.vol);
    *g.vol = n ? (unsigned int*)calloc(n, sizeof(unsigned int)) : nullptr;
    if (o.get("Tag") && *g.vol) {
        unsigned int tag = (unsigned int)tag_of(o);
        for (size_t i = 0; i < n; i++) (*g.vol)[i] = tag;
    }
}

static void shape_sphere(const Json& o, Grid& g) {
    const Json* c = o.get("O");
    if (!c || c->size() != 3) throw Error(-1, "A Sphere command misses O field");
    const Json* r = o.get("R");
    if (!r) throw Error(-1, "A Sphere command misses R field");
    float O[3] = {(float)num_at(*c, 0), (float)num_at(*c, 1), (float)num_at(*c, 2)};
    float R = (float)r->as_num(), R2 = R * R;
    unsigned int tag = (unsigned int)tag_of(o);
    for (int k = 0; k < (int)g.dim->z; k++) {
        float dz = (k + 0.5f) - O[2];
        for (int j = 0; j < (int)g.dim->y; j++) {
            float dy = (j + 0.5f) - O[1];
            for (int i = 0; i < (int)g.dim->x; i++) {
                float dx = (i + 0.5f) - O[0];
                if (dx * dx + dy * dy + dz * dz <= R2) (*g.vol)[vidx(g, i, j, k)] = tag;
            }
        }
    }
}

static void shape_subgrid(const Json& o, Grid& g) {
    const Json* c = o.get("O");
    if (!c || c->size() != 3) throw Error(-1, "A Subgrid command misses O field");
    const Json* sz = o.get("Size");
    if (!sz || sz->size() != 3) throw Error(-1, "A Subgrid command misses Size field");
    int O[3] = {(int)num_at(*c, 0) - 1 - (int)g.ox, (int)num_at(*c, 1) - 1 - (int)g.oy, (int)num_at(*c, 2) - 1 - (int)g.oz};
    int S[3] = {(int)num_at(*sz, 0), (int)num_at(*sz, 1), (int)num_at(*sz, 2)};
    unsigned int tag = (unsigned int)tag_of(o);
    for (int k = 0; k < (int)g.dim->z; k++) {
        if (k < O[2] || k > O[2] + S[2]) continue;
        for (int j = 0; j < (int)g.dim->y; j++) {
            if (j < O[1] || j > O[1] + S[1]) continue;
            for (int i = 0; i < (int)g.dim->x; i++) {
                if (i < O[0] || i > O[0] + S[0]) continue;
                (*g.vol)[vidx(g, i, j, k)] = tag;
            }
        }
    }
}

static void shape_box(const Json& o, Grid& g) {
    const Json* c = o.get("O");
    if (!c || c->size() != 3) throw Error(-1, "A Box command misses O field");
    const Json* sz = o.get("Size");
    if (!sz || sz->size() != 3) throw Error(-1, "A Box command misses Size field");
    float O[3] = {(float)(num_at(*c, 0) - g.ox), (float)(num_at(*c, 1) - g.oy), (float)(num_at(*c, 2) - g.oz)};
    float S[3] = {(float)num_at(*sz, 0), (float)num_at(*sz, 1), (float)num_at(*sz, 2)};
    unsigned int tag = (unsigned int)tag_of(o);
    for (int k = 0; k < (int)g.dim->z; k++) {
        float dz = k + 0.5f;
        if (dz < O[2] || dz > O[2] + S[2]) continue;
        for (int j = 0; j < (int)g.dim->y; j++) {
            float dy = j + 0.5f;
            if (dy < O[1] || dy > O[1] + S[1]) continue;
            for (int i = 0; i < (int)g.dim->x; i++) {
                float dx = i + 0.5f;
                if (dx < O[0] || dx > O[0] + S[0]) continue;
                (*g.vol)[vidx(g, i, j, k)] = tag;
            }
        }
    }
}

static void shape_cylinder(const Json& o, Grid& g) {
    const Json* a = o.get("C0");
    const Json* b = o.get("C1");
    if (!a || a->size() != 3) throw Error(-1, "A Cylinder command misses C0 field");
    if (!b || b->size() != 3) throw Error(-1, "A Cylinder command misses C1 field");
    float C0[3] = {(float)(num_at(*a, 0) - g.ox), (float)(num_at(*a, 1) - g.oy), (float)(num_at(*a, 2) - g.oz)};
    float C1[3] = {(float)(num_at(*b, 0) - g.ox), (float)(num_at(*b, 1) - g.oy), (float)(num_at(*b, 2) - g.oz)};
    float v[3] = {C1[0] - C0[0], C1[1] - C0[1], C1[2] - C0[2]};
    float d0 = (float)sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    if (d0 == 0.f) throw Error(-1, "Coincident end points in the definition of Cylinder command");
    v[0] /= d0; v[1] /= d0; v[2] /= d0;
    const Json* r = o.get("R");
    if (!r) throw Error(-1, "A Cylinder command misses R field");
    float R = (float)r->as_num(), R2 = R * R;
    unsigned int tag = (unsigned int)tag_of(o);
    for (int k = 0; k < (int)g.dim->z; k++) {
        float dz = (k + 0.5f) - C0[2];
        for (int j = 0; j < (int)g.dim->y; j++) {
            float dy = (j + 0.5f) - C0[1];
            for (int i = 0; i < (int)g.dim->x; i++) {
                float dx = (i + 0.5f) - C0[0];
                float d = v[0] * dx + v[1] * dy + v[2] * dz;
                if (d > d0 || d < 0.f) continue;
                d = dx * dx + dy * dy + dz * dz - d * d;
                if (d <= R2) (*g.vol)[vidx(g, i, j, k)] = tag;
            }
        }
    }
}

static void shape_layers(const std::string& name, const Json& o, Grid& g) {
    int dir = name[0] - 'X';
    if (!o.is_arr()) throw Error(-1, "A " + name + " object must be an array");
    std::vector<const Json*> items;
    if (o.size() == 3 && !o.arr[0].is_arr())
        items.push_back(&o);
    else
        for (const auto& x : o.arr) items.push_back(&x);
    const unsigned int dims[3] = {g.dim->x, g.dim->y, g.dim->z};
    for (const Json* it : items) {
        if (it->size() != 3) throw Error(-1, "The " + name + " must contain integer triplets");
        int lo = std::max((int)num_at(*it, 0), 1) - 1;
        int hi = std::min((int)num_at(*it, 1), (int)dims[dir]);
        int tag = (int)num_at(*it, 2);
        if (hi < lo) std::swap(lo, hi);
        for (int k = 0; k < (int)g.dim->z; k++)
            for (int j = 0; j < (int)g.dim->y; j++)
                for (int i = 0; i < (int)g.dim->x; i++) {
                    int c = dir == 0 ? i : (dir == 1 ? j : k);
                    if (c >= lo && c < hi) (*g.vol)[vidx(g, i, j, k)] = (unsigned int)tag;
                }
    }
}

static void shape_slabs(const std::string& name, const Json& o, Grid& g) {
    int dir = name[0] - 'X';
    const Json* b = o.get("Bound");
    if (!b || !b->is_arr()) throw Error(-1, "A " + name + " command misses Bound field or not an array");
    std::vector<const Json*> items;
    if (b->size() == 2 && !b->arr[0].is_arr())
        items.push_back(b);
    else
        for (const auto& x : b->arr) items.push_back(&x);
    const unsigned int dims[3] = {g.dim->x, g.dim->y, g.dim->z};
    unsigned int tag = (unsigned int)tag_of(o);
    for (const Json* it : items) {
        if (it->size() != 2) throw Error(-1, "The Bound field must contain number pairs");
        float lo = (float)std::max(num_at(*it, 0) - 0.5, 0.0);
        float hi = (float)std::min(num_at(*it, 1), (double)dims[dir]);
        if (hi < lo) std::swap(lo, hi);
        for (int k = 0; k < (int)g.dim->z; k++)
            for (int j = 0; j < (int)g.dim->y; j++)
                for (int i = 0; i < (int)g.dim->x; i++) {
                    int c = dir == 0 ? i : (dir == 1 ? j : k);
                    if (c >= (int)lo && c < (int)hi) (*g.vol)[vidx(g, i, j, k)] = tag;
                }
    }
}

static void shape_upperspace(const Json& o, Grid& g) {
    const Json* c = o.get("Coef");
    if (!c || c->size() != 4) throw Error(-1, "An UpperSpace command misses Coef field");
    float C[4] = {(float)num_at(*c, 0), (float)num_at(*c, 1), (float)num_at(*c, 2), (float)num_at(*c, 3)};
    unsigned int tag = (unsigned int)tag_of(o);
    for (int k = 0; k < (int)g.dim->z; k++)
        for (int j = 0; j < (int)g.dim->y; j++)
            for (int i = 0; i < (int)g.dim->x; i++)
                if (C[0] * (i + 0.5f) + C[1] * (j + 0.5f) + C[2] * (k + 0.5f) > C[3]) (*g.vol)[vidx(g, i, j, k)] = tag;
}

/* src/mcx_shapes.c:163-235: apply each single-key shape object in order */
static void rasterize_shapes(const Json& shapes, Grid& g) {
    size_t n = (size_t)g.dim->x * g.dim->y * g.dim->z;
    if (n > 0) {
        if (*g.vol)
            *g.vol = (unsigned int*)realloc(*g.vol, n * sizeof(unsigned int));
        else
            *g.vol = (unsigned int*)calloc(n, sizeof(unsigned int));
    }
    if (!shapes.is_arr()) throw Error(-1, "the Shapes section must be an array");
    int count = 1;
    for (const Json& item : shapes.arr) {
        if (!item.is_obj() || item.obj.empty()) break;
        const std::string& k = item.obj[0].first;
        const Json& o = item.obj[0].second;
        if (k == "Name") {
        } else if (k == "Origin") {
            if (o.size() != 3) throw Error(-1, "An Origin record does not contain a triplet");
            g.ox = (float)num_at(o, 0);
            g.oy = (float)num_at(o, 1);
            g.oz = (float)num_at(o, 2);
        } else if (k == "Grid") {
            shape_grid(o, g);
        } else if (k == "Sphere") {
            shape_sphere(o, g);
        } else if (k == "Box") {
            shape_box(o, g);
        } else if (k == "Subgrid") {
            shape_subgrid(o, g);
        } else if (k == "Cylinder") {
            shape_cylinder(o, g);
        } else if (k == "XLayers" || k == "YLayers" || k == "ZLayers") {
            shape_layers(k, o, g);
        } else if (k == "XSlabs" || k == "YSlabs" || k == "ZSlabs") {
            shape_slabs(k, o, g);
        } else if (k == "UpperSpace") {
            shape_upperspace(o, g);
        } else {
            throw Error(-(count + 100), "The #" + std::to_string(count) + " element in the Shapes section has an undefined tag " + k);
        }
        count++;
    }
}

/* ---- JSON -> Config (src/mcx_utils.c:2257-3210) ---------------------- */

/* src/mcx_utils.c:3510-3600 (mcx_loadvolume): labels widen to 32 bits;
 * float media are encoded into property words; the other non-label
 * formats are stored as read */
static void load_volume_file(Config* cfg, const std::string& path) {
    std::ifstream f(path, std::ios::binary);
    if (!f) throw Error(-1, "can not open the volume file " + path);
    const size_t n = (size_t)cfg->dim.x * cfg->dim.y * cfg->dim.z;
    const unsigned int fmt = cfg->mediabyte;
    const size_t per = fmt == MCX_MEDIA_ASGN_F2H ? 16 : (fmt == MCX_MEDIA_AS_F2H || fmt == MCX_MEDIA_2LABEL_SPLIT) ? 8 : std::min(fmt, 4u);
    const size_t words = (fmt == MCX_MEDIA_ASGN_F2H || fmt == MCX_MEDIA_2LABEL_SPLIT) ? 2 * n : n;
    std::vector<unsigned char> raw(n * per);
    f.read((char*)raw.data(), (std::streamsize)raw.size());
    if ((size_t)f.gcount() != raw.size()) throw Error(-1, "volume file is smaller than Domain.Dim");
    free(cfg->vol);
    cfg->vol = (unsigned int*)calloc(words, sizeof(unsigned int));
    const float* fv = (const float*)raw.data();
    for (size_t i = 0; i < n; i++) {
        if (fmt == 1) {
            cfg->vol[i] = raw[i];
        } else if (fmt == 2) {
            cfg->vol[i] = ((const unsigned short*)raw.data())[i];
        } else if (fmt == MCX_MEDIA_MUA_FLOAT) {
            cfg->vol[i] = encode_mua_float(fv[i], cfg->unitinmm, true);
        } else if (fmt == MCX_MEDIA_AS_F2H || fmt == MCX_MEDIA_ASGN_F2H) {
            const size_t k = fmt == MCX_MEDIA_ASGN_F2H ? 4 * i : 2 * i;
            const float mua = fv[k] * cfg->unitinmm, mus = fv[k + 1] * cfg->unitinmm;
            cfg->vol[i] = (mua != mua || mus != mus) ? 0u : encode_half2(mua, mus);
            if (fmt == MCX_MEDIA_ASGN_F2H) cfg->vol[i + n] = encode_half2(fv[k + 2], fv[k + 3]);
        } else if (fmt == MCX_MEDIA_2LABEL_SPLIT) {
            memcpy(cfg->vol + 2 * i, raw.data() + 8 * i, 8);
        } else {
            memcpy(cfg->vol + i, raw.data() + 4 * i, 4);
        }
    }
}

static void loadjson(const Json& root, Config* cfg, std::string* volfile) {
    const Json* Domain = root.get("Domain");
    const Json* Optode = root.get("Optode");
    const Json* Session = root.get("Session");
    const Json* Forward = root.get("Forward");
    const Json* Shapes = root.get("Shapes");
    if (Domain) {
        if (const Json* v = Domain->get("VolumeFile")) {
            if (v->is_str() && !v->str.empty()) *volfile = v->str;
        }
        if (const Json* v = Domain->get("MediaFormat")) {
            int k = key_lookup(v->str, kMediaFormats);
            if (k < 0) throw Error(-1, "Unsupported media format.");
            cfg->mediabyte = (unsigned int)kMediaFormatIds[k];
        }
        if (const Json* v = Domain->get("LengthUnit")) cfg->unitinmm = (float)v->as_num(1.0);
        if (const Json* meds = Domain->get("Media")) {
            if (meds->is_arr() && meds->size()) {
                cfg->medianum = (unsigned int)meds->size();
                free(cfg->prop);
                cfg->prop = (Medium*)malloc(sizeof(Medium) * cfg->medianum);
                for (unsigned int i = 0; i < cfg->medianum; i++) {
                    const Json& m = meds->arr[i];
                    if (m.is_obj()) {
                        const Json *a = m.get("mua"), *s = m.get("mus"), *g = m.get("g"), *n = m.get("n");
                        if (!a) throw Error(-1, "You must specify absorption coeff, default in 1/mm");
                        if (!s) throw Error(-1, "You must specify scattering coeff, default in 1/mm");
                        if (!g) throw Error(-1, "You must specify anisotropy [0-1]");
                        if (!n) throw Error(-1, "You must specify refractive index");
                        cfg->prop[i] = {(float)a->as_num(), (float)s->as_num(), (float)g->as_num(), (float)n->as_num()};
                    } else if (m.is_arr()) {
                        cfg->prop[i] = {(float)num_at(m, 0), (float)num_at(m, 1), (float)num_at(m, 2), (float)num_at(m, 3)};
                    } else {
                        throw Error(-1, "Session.Media must be either an array of objects or array of 4-elem numerical arrays");
                    }
                }
            }
        }
        if (Domain->get("MieScatter")) throw Error(-1, "polarized photon simulation is not supported by this engine");
        if (const Json* v = Domain->get("Dim")) {
            if (v->size() >= 3) {
                cfg->dim.x = (unsigned)num_at(*v, 0);
                cfg->dim.y = (unsigned)num_at(*v, 1);
                cfg->dim.z = (unsigned)num_at(*v, 2);
            }
        } else if (!Shapes) {
            throw Error(-1, "You must specify the dimension of the volume");
        }
        if (const Json* v = Domain->get("Step")) {
            if (v->size() >= 3) {
                cfg->steps.x = (float)num_at(*v, 0);
                cfg->steps.y = (float)num_at(*v, 1);
                cfg->steps.z = (float)num_at(*v, 2);
            }
            if (cfg->steps.x != cfg->steps.y || cfg->steps.y != cfg->steps.z)
                throw Error(-1, "MCX currently does not support anisotropic voxels");
        }
        if (cfg->steps.x != 1.f && cfg->unitinmm == 1.f) cfg->unitinmm = cfg->steps.x;
        if (const Json* v = Domain->get("OriginType")) {
            if (cfg->issrcfrom0 == 0) cfg->issrcfrom0 = (char)v->as_num();
        }
    }
    if (Optode) {
        if (const Json* src = Optode->get("Source")) {
            if (const Json* v = src->get("Pos")) {
                if (v->size() && v->arr[0].is_arr()) throw Error(-1, "multiple sources are not supported by this engine");
                cfg->srcpos.x = (float)num_at(*v, 0);
                cfg->srcpos.y = (float)num_at(*v, 1);
                cfg->srcpos.z = (float)num_at(*v, 2);
                if (v->size() > 3) cfg->srcpos.w = (float)num_at(*v, 3);
            }
            if (const Json* v = src->get("Weight")) cfg->srcpos.w = (float)v->as_num();
            if (const Json* v = src->get("Dir")) {
                cfg->srcdir.x = (float)num_at(*v, 0);
                cfg->srcdir.y = (float)num_at(*v, 1);
                cfg->srcdir.z = (float)num_at(*v, 2);
                if (v->size() > 3) {
                    const Json& w = v->arr[3];
                    if (w.is_str()) {
                        if (w.str == "_NaN_") cfg->srcdir.w = NAN;
                        else if (w.str == "-_Inf_") cfg->srcdir.w = -INFINITY;
                        else if (w.str == "_Inf_") cfg->srcdir.w = INFINITY;
                    } else {
                        cfg->srcdir.w = (float)w.as_num();
                    }
                }
            }
            if (const Json* v = src->get("Type")) {
                int k = key_lookup(v->str, kSrcTypes);
                if (k < 0) throw Error(-1, "the specified source type is not supported");
                cfg->srctype = (char)k;
            }
            if (const Json* v = src->get("Param1")) {
                float* p = &cfg->srcparam1.x;
                for (size_t i = 0; i < std::min<size_t>(4, v->size()); i++) p[i] = (float)num_at(*v, i);
            }
            if (const Json* v = src->get("Param2")) {
                float* p = &cfg->srcparam2.x;
                for (size_t i = 0; i < std::min<size_t>(4, v->size()); i++) p[i] = (float)num_at(*v, i);
            }
            if (const Json* v = src->get("Frequency")) cfg->omega = (float)v->as_num();
            if (const Json* v = src->get("WaveLength")) cfg->lambda = (float)v->as_num();
            if (const Json* v = src->get("SrcNum")) cfg->srcnum = (unsigned int)v->as_num();
            if (src->get("Pattern")) throw Error(-1, "pattern sources are not supported by this engine");
            if (!cfg->issrcfrom0) {
                cfg->srcpos.x--;
                cfg->srcpos.y--;
                cfg->srcpos.z--;
            }
        }
        if (const Json* dets = Optode->get("Detector")) {
            if (dets->is_arr() && dets->size()) {
                cfg->detnum = (unsigned int)dets->size();
                free(cfg->detpos);
                cfg->detpos = (mcx_float4*)malloc(sizeof(mcx_float4) * cfg->detnum);
                for (unsigned int i = 0; i < cfg->detnum; i++) {
                    const Json& d = dets->arr[i];
                    const Json* pos = d.get("Pos");
                    const Json* r = d.get("R");
                    cfg->detpos[i] = {0.f, 0.f, 0.f, cfg->detradius};
                    if (pos) {
                        cfg->detpos[i].x = (float)num_at(*pos, 0);
                        cfg->detpos[i].y = (float)num_at(*pos, 1);
                        cfg->detpos[i].z = (float)num_at(*pos, 2);
                    }
                    if (r) cfg->detpos[i].w = (float)r->as_num();
                    if (!cfg->issrcfrom0) {
                        cfg->detpos[i].x--;
                        cfg->detpos[i].y--;
                        cfg->detpos[i].z--;
                    }
                }
            }
        }
    }
    if (cfg->medianum + cfg->detnum > MCX_MAX_PROP_AND_DETECTORS)
        throw Error(-4, "input media types plus detector number plus additional sources exceeds the maximum total (4000)");
    if (Session) {
        if (const Json* v = Session->get("RNGSeed")) cfg->seed = (int)(long long)v->as_num(-1);
        else cfg->seed = -1;
        if (const Json* v = Session->get("Photons")) cfg->nphoton = (size_t)v->as_num();
        if (const Json* v = Session->get("MaxDetPhoton")) cfg->maxdetphoton = (unsigned int)v->as_num();
        if (const Json* v = Session->get("ID")) strncpy(cfg->session, v->str.c_str(), MCX_MAX_SESSION_LENGTH - 1);
        if (const Json* v = Session->get("RootPath")) strncpy(cfg->rootpath, v->str.c_str(), MCX_MAX_PATH_LENGTH - 1);
        if (const Json* v = Session->get("BCFlags")) {
            memset(cfg->bc, 0, sizeof(cfg->bc));
            strncpy(cfg->bc, v->str.c_str(), 12);
        }
        if (const Json* v = Session->get("DoMismatch")) cfg->isreflect = (char)v->as_num();
        if (const Json* v = Session->get("DoSaveVolume")) cfg->issave2pt = (char)v->as_num();
        if (const Json* v = Session->get("DoNormalize")) cfg->isnormalized = (char)v->as_num();
        if (const Json* v = Session->get("DoPartialPath")) cfg->issavedet = (char)v->as_num();
        if (const Json* v = Session->get("DoSaveRef")) cfg->issaveref = (char)v->as_num();
        if (const Json* v = Session->get("DoSaveExit")) cfg->issaveexit = (char)v->as_num();
        if (const Json* v = Session->get("DoSaveSeed")) cfg->issaveseed = (char)v->as_num();
        if (const Json* v = Session->get("DoAutoThread")) cfg->autopilot = (char)v->as_num();
        if (const Json* v = Session->get("DoDCS")) cfg->ismomentum = (char)v->as_num();
        if (const Json* v = Session->get("DoSpecular")) cfg->isspecular = (char)v->as_num();
        if (const Json* v = Session->get("DebugFlag"))
            cfg->debuglevel = v->is_str() ? parse_flags(v->str, kDebugFlags) : (unsigned int)v->as_num();
        if (const Json* v = Session->get("SaveDataMask"))
            cfg->savedetflag = v->is_str() ? parse_flags(v->str, kSaveFlags) : (unsigned int)v->as_num();
        if (const Json* v = Session->get("OutputFormat")) {
            int k = key_lookup(v->str, kOutFormats);
            if (k < 0) throw Error(-2, "the specified output format is not recognized");
            cfg->outputformat = (char)k;
        }
        if (const Json* v = Session->get("OutputType")) {
            const char* p = v->str.empty() ? nullptr : strchr(kOutTypes, tolower((unsigned char)v->str[0]));
            if (!p) throw Error(-2, "the specified output data type is not recognized");
            cfg->outputtype = (char)(p - kOutTypes);
        }
        if (const Json* v = Session->get("MinEnergy")) cfg->minenergy = (float)v->as_num();
    }
    if (Forward) {
        const Json* t0 = Forward->get("T0");
        const Json* t1 = Forward->get("T1");
        const Json* dt = Forward->get("Dt");
        cfg->tstart = t0 ? (float)t0->as_num() : 0.f;
        cfg->tend = t1 ? (float)t1->as_num() : 0.f;
        cfg->tstep = dt ? (float)dt->as_num() : 0.f;
        if (cfg->tstart > cfg->tend || cfg->tstep == 0.f) throw Error(-9, "incorrect time gate settings");
        unsigned int gates = (unsigned int)((cfg->tend - cfg->tstart) / cfg->tstep + 0.5);
        if (cfg->maxgate == 0 || cfg->maxgate > gates) cfg->maxgate = gates;
    }
    if (Shapes) {
        if (cfg->mediabyte > 4) throw Error(-10, "rasterization of shapes must be used with label-based mediatype");
        Grid g{&cfg->vol, &cfg->dim, 1.f, 1.f, 1.f};
        if (cfg->issrcfrom0) g.ox = g.oy = g.oz = 0.f;
        if (Shapes->is_obj() && Shapes->get("_ArraySize_"))
            throw Error(-1, "JData-encoded volumes in Shapes are not supported by this engine");
        rasterize_shapes(*Shapes, g);
    }
    /* src/mcx_utils.c:3203-3207: the detected-photon file header */
    cfg->his.maxmedia = cfg->medianum - 1;
    cfg->his.detnum = cfg->detnum;
    cfg->his.srcnum = cfg->srcnum;
    cfg->his.savedetflag = cfg->savedetflag;
    cfg->his.totalsource = 1;
}

static std::string read_text(const char* fname) {
    std::ifstream f(fname, std::ios::binary);
    if (!f) throw Error(-2, std::string("can not open input file ") + fname);
    std::stringstream ss;
    ss << f.rdbuf();
    return ss.str();
}

/* ---- base64 + zlib (JData _ArrayZipData_) ----------------------------- */

static std::string base64(const unsigned char* d, size_t n) {
    static const char* t = "ABCDEFGHIJKLMNOPQRSTUVWXYZabcdefghijklmnopqrstuvwxyz0123456789+/";
    std::string o;
    o.reserve((n + 2) / 3 * 4);
    size_t i = 0;
    for (; i + 2 < n; i += 3) {
        unsigned v = (d[i] << 16) | (d[i + 1] << 8) | d[i + 2];
        o += t[v >> 18]; o += t[(v >> 12) & 63]; o += t[(v >> 6) & 63]; o += t[v & 63];
    }
    if (i + 1 == n) {
        unsigned v = d[i] << 16;
        o += t[v >> 18]; o += t[(v >> 12) & 63]; o += "==";
    } else if (i + 2 == n) {
        unsigned v = (d[i] << 16) | (d[i + 1] << 8);
        o += t[v >> 18]; o += t[(v >> 12) & 63]; o += t[(v >> 6) & 63]; o += '=';
    }
    return o;
}

static Json jdata_encode(const void* data, size_t nbytes, const std::vector<unsigned int>& dims, const char* type, bool colmajor) {
    uLongf clen = compressBound((uLong)nbytes);
    std::vector<unsigned char> buf(clen);
    if (compress2(buf.data(), &clen, (const Bytef*)data, (uLong)nbytes, Z_DEFAULT_COMPRESSION) != Z_OK)
        throw Error(-1, "error when compressing data");
    size_t count = 1;
    for (unsigned d : dims) count *= d;
    Json o = Json::object();
    o.set("_ArrayType_", Json::string(type));
    Json& sz = o.set("_ArraySize_", Json::array());
    for (unsigned d : dims) sz.push(Json::number(d));
    if (colmajor) o.set("_ArrayOrder_", Json::string("c"));
    o.set("_ArrayZipType_", Json::string("zlib"));
    o.set("_ArrayZipSize_", Json::number((double)count));
    o.set("_ArrayZipData_", Json::string(base64(buf.data(), clen)));
    return o;
}

static std::string out_name(const Config* cfg) {
    if (cfg->rootpath[0]) return std::string(cfg->rootpath) + "/" + cfg->session;
    return std::string(cfg->session);
}

static void write_text(const std::string& fname, const std::string& s) {
    FILE* fp = fopen(fname.c_str(), "wt");
    if (!fp) throw Error(-2, "can not save data to disk");
    fprintf(fp, "%s\n", s.c_str());
    fclose(fp);
}

/* src/mcx_utils.c:791-905 */
static void save_jnii(const float* vol, const std::vector<unsigned int>& dims, const std::vector<float>& voxelsize,
                      const std::string& name, const Config* cfg) {
    Json root = Json::object();
    Json& info = root.set("_DataInfo_", Json::object());
    info.set("JNIFTIVersion", Json::string("0.5"));
    info.set("Comment", Json::string("Created by MCX (MI355X engine)"));
    info.set("AnnotationFormat", Json::string("https://neurojson.org/jnifti/draft1"));
    info.set("SerialFormat", Json::string("https://json.org"));
    Json& hdr = root.set("NIFTIHeader", Json::object());
    hdr.set("NIIHeaderSize", Json::number(348));
    Json& d = hdr.set("Dim", Json::array());
    for (unsigned v : dims) d.push(Json::number(v));
    hdr.set("Param1", Json::number(0));
    hdr.set("Param2", Json::number(0));
    hdr.set("Param3", Json::number(0));
    hdr.set("Intent", Json::number(0));
    hdr.set("DataType", Json::string("single"));
    hdr.set("BitDepth", Json::number(32));
    hdr.set("FirstSliceID", Json::number(0));
    Json& vs = hdr.set("VoxelSize", Json::array());
    for (float v : voxelsize) vs.push(Json::number(v));
    Json& ori = hdr.set("Orientation", Json::object());
    ori.set("x", Json::string("r"));
    ori.set("y", Json::string("a"));
    ori.set("z", Json::string("s"));
    hdr.set("ScaleSlope", Json::number(1));
    hdr.set("ScaleOffset", Json::number(0));
    hdr.set("LastSliceID", Json::number(cfg->maxgate));
    hdr.set("SliceType", Json::number(1));
    Json& unit = hdr.set("Unit", Json::object());
    unit.set("L", Json::string("mm"));
    unit.set("T", Json::string("s"));
    hdr.set("MaxIntensity", Json::number(1));
    hdr.set("MinIntensity", Json::number(0));
    hdr.set("SliceTime", Json::number(0));
    hdr.set("TimeOffset", Json::number(0));
    static const char* desc[] = {"MCX volumetric output: Fluence rate (W/mm^2)", "MCX volumetric output: Fluence (J/mm^2)",
                                 "MCX volumetric output: Voxel-wise energy deposit (J)", "MCX volumetric output: Jacobian for mua (J/mm)",
                                 "MCX volumetric output: Scattering count", "MCX volumetric output: Partial momentum transfer"};
    hdr.set("Description", Json::string(cfg->outputtype >= 0 && cfg->outputtype < 6 ? desc[(int)cfg->outputtype] : "MCX volumetric output"));
    hdr.set("Name", Json::string(cfg->session));
    hdr.set("NIIFormat", Json::string("jnifti"));
    size_t n = 1;
    for (unsigned v : dims) n *= v;
    root.set("NIFTIData", jdata_encode(vol, n * sizeof(float), dims, "single", true));
    write_text(name + ".jnii", root.dump());
}

/* src/mcx_utils.c:1029-1201 */
static void save_jdet(const float* ppath, const void* seeds, unsigned int count, const Config* cfg) {
    Json root = Json::object();
    Json& obj = root.set("MCXData", Json::object());
    Json& hdr = obj.set("Info", Json::object());
    hdr.set("Version", Json::number(cfg->his.version));
    hdr.set("MediaNum", Json::number(cfg->his.maxmedia));
    hdr.set("DetNum", Json::number(cfg->his.detnum));
    hdr.set("ColumnNum", Json::number(cfg->his.colcount));
    hdr.set("TotalPhoton", Json::number(cfg->his.totalphoton));
    hdr.set("DetectedPhoton", Json::number(count));
    hdr.set("SavedPhoton", Json::number(cfg->his.savedphoton));
    hdr.set("LengthUnit", Json::number(cfg->his.unitinmm));
    hdr.set("SeedByte", Json::number(cfg->his.seedbyte));
    hdr.set("Normalizer", Json::number(cfg->his.normalizer));
    hdr.set("Repeat", Json::number(cfg->his.respin));
    hdr.set("SrcNum", Json::number(cfg->his.srcnum));
    hdr.set("SaveDetFlag", Json::number(cfg->his.savedetflag));
    hdr.set("TotalSource", Json::number(cfg->his.totalsource));
    Json& media = hdr.set("Media", Json::array());
    for (unsigned int i = 0; i < cfg->medianum; i++) {
        Json m = Json::object();
        m.set("mua", Json::number(cfg->prop[i].mua / cfg->unitinmm));
        m.set("mus", Json::number(cfg->prop[i].mus / cfg->unitinmm));
        m.set("g", Json::number(cfg->prop[i].g));
        m.set("n", Json::number(cfg->prop[i].n));
        media.push(m);
    }
    const unsigned int colnum[] = {1, cfg->his.maxmedia, cfg->his.maxmedia, cfg->his.maxmedia, 3, 3, 1, 4};
    const bool isint[] = {true, true, false, false, false, false, false, false};
    const char* dname[] = {"detid", "nscat", "ppath", "mom", "p", "v", "w0", "s"};
    Json& dat = obj.set("PhotonData", Json::object());
    unsigned int col = 0;
    for (int id = 0; id < 8; id++) {
        if (!((cfg->savedetflag >> id) & 1)) continue;
        std::vector<unsigned int> dims = {count, colnum[id]};
        std::vector<unsigned int> ib;
        std::vector<float> fb;
        for (unsigned int i = 0; i < count; i++)
            for (unsigned int j = 0; j < colnum[id]; j++) {
                float v = ppath[(size_t)i * cfg->his.colcount + col + j];
                if (isint[id])
                    ib.push_back((unsigned int)v);
                else
                    fb.push_back(v);
            }
        if (isint[id])
            dat.set(dname[id], jdata_encode(ib.data(), ib.size() * 4, dims, "uint32", false));
        else
            dat.set(dname[id], jdata_encode(fb.data(), fb.size() * 4, dims, "single", false));
        col += colnum[id];
    }
    if (cfg->issaveseed && seeds) {
        std::vector<unsigned int> dims = {count, cfg->his.seedbyte};
        dat.set("seed", jdata_encode(seeds, (size_t)count * cfg->his.seedbyte, dims, "uint8", false));
    }
    write_text(out_name(cfg) + "_detp.jdat", root.dump());
}

/* pmcx 'shapes' string (src/mcx_utils.c:3822-3837): a JSON object with a
 * Shapes member, or a bare Shapes array, rasterized over cfg->vol */
void apply_shapes(Config* cfg, const char* text) {
    if (!text || !strchr(text, ':')) return;
    if (cfg->mediabyte > 4) throw Error(-6, "rasterization of shapes must be used with label-based mediatype");
    Json j = Json::parse(text);
    const Json* shapes = j.is_obj() ? j.get("Shapes") : &j;
    if (!shapes) return;
    Grid g{&cfg->vol, &cfg->dim, 1.f, 1.f, 1.f};
    if (cfg->issrcfrom0) g.ox = g.oy = g.oz = 0.f;
    rasterize_shapes(*shapes, g);
}

}  // namespace mcx

using namespace mcx;

extern "C" {

void mcx_readconfig(char* fname, Config* cfg) {
    EntryGuard g;
    try {
        std::string text;
        if (fname[0] == '{')
            text = fname;
        else
            text = read_text(fname);
        Json root = Json::parse(text);
        std::string volfile;
        loadjson(root, cfg, &volfile);
        if (!cfg->session[0]) {
            std::string s = fname[0] == '{' ? "default" : fname;
            size_t p = s.find_last_of('/');
            if (p != std::string::npos) s = s.substr(p + 1);
            p = s.rfind(".json");
            if (p != std::string::npos) s = s.substr(0, p);
            strncpy(cfg->session, s.c_str(), MCX_MAX_SESSION_LENGTH - 1);
        }
        if (!volfile.empty()) {
            std::string path = volfile;
            if (cfg->rootpath[0] && volfile[0] != '/') path = std::string(cfg->rootpath) + "/" + volfile;
            load_volume_file(cfg, path);
        }
        if (!cfg->vol) throw Error(-4, "one must specify a binary volume file in order to run the simulation");
        mcx_preprocess(cfg);
    } catch (const Error& e) {
        handle_error(e, cfg);
    }
}

int mcx_run_from_json(char* jsonstr) {
    Config cfg;
    GPUInfo* gpuinfo = NULL;
    mcx_initcfg(&cfg);
    mcx_readconfig(jsonstr, &cfg);
    int n = mcx_list_gpu(&cfg, &gpuinfo);
    if (!n) {
        mcx_clearcfg(&cfg);
        return -1;
    }
    mcx_run_simulation(&cfg, gpuinfo);
    mcx_cleargpuinfo(&gpuinfo);
    mcx_clearcfg(&cfg);
    return 0;
}

void mcx_savedata(float* dat, size_t len, Config* cfg) {
    EntryGuard g;
    try {
        std::string name = out_name(cfg);
        /* src/mcx_utils.c:929-951: per-detector replay blocks and the RF
         * imaginary plane form a sixth dimension; RF data is twice len */
        const bool replay = cfg->seed == MCX_SEED_FROM_FILE;
        const bool rf = replay && (cfg->outputtype == otRF || cfg->outputtype == otRFmus);
        unsigned int d5 = 1;
        if (replay && cfg->replaydet == -1 && cfg->detnum > 1) d5 *= cfg->detnum;
        if (rf) d5 *= 2;
        if (cfg->outputformat == ofJNifti || cfg->outputformat == ofBJNifti) {
            std::vector<unsigned int> dims = {cfg->dim.x, cfg->dim.y, cfg->dim.z, cfg->maxgate, cfg->srcnum};
            std::vector<float> vs = {cfg->steps.x, cfg->steps.y, cfg->steps.z, cfg->tstep, 1.f};
            if (d5 > 1) {
                dims.push_back(d5);
                vs.push_back(1.f);
            }
            save_jnii(dat, dims, vs, name, cfg);
            return;
        }
        if (rf) len *= 2;
        if (cfg->outputformat == ofNifti || cfg->outputformat == ofAnalyze || cfg->outputformat == ofUBJSON || cfg->outputformat == ofTX3)
            throw Error(-2, "this output format is not supported by this engine; use jnii or mc2");
        std::string fname = name + ".mc2";
        FILE* fp = fopen(fname.c_str(), "wb");
        if (!fp) throw Error(-2, "can not save data to disk");
        fwrite(dat, sizeof(float), len, fp);
        fclose(fp);
    } catch (const Error& e) {
        handle_error(e, cfg);
    }
}

void mcx_savedetphoton(float* ppath, void* seeds, int count, int seedbyte, Config* cfg) {
    (void)seedbyte;
    EntryGuard g;
    try {
        if (cfg->outputformat == ofJNifti || cfg->outputformat == ofBJNifti) {
            save_jdet(ppath, seeds, (unsigned int)count, cfg);
            return;
        }
        std::string fname = out_name(cfg) + ".mch";
        FILE* fp = fopen(fname.c_str(), "wb");
        if (!fp) throw Error(-2, "can not save data to disk");
        fwrite(&cfg->his, sizeof(History), 1, fp);
        fwrite(ppath, sizeof(float), (size_t)count * cfg->his.colcount, fp);
        if (cfg->issaveseed && seeds) fwrite(seeds, cfg->his.seedbyte, (size_t)count, fp);
        fclose(fp);
    } catch (const Error& e) {
        handle_error(e, cfg);
    }
}

}  // extern "C"
