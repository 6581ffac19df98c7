/*
 * mcx_internal.h -- shared host-side helpers of the library (not ABI).
 */
#ifndef MCX_AMD_INTERNAL_H
#define MCX_AMD_INTERNAL_H

#include <stddef.h>
#include <stdint.h>

#include <stdexcept>
#include <string>

#include "../../include/mcx_abi.h"

namespace mcx {

/* Error raised inside the library; entry points translate it into the
 * reference's convention (print + exit for standalone callers, recorded
 * error for embedded ones). */
struct Error : public std::exception {
    int id;
    std::string msg;
    Error(int i, std::string m) : id(i), msg(std::move(m)) {}
    const char* what() const noexcept override { return msg.c_str(); }
};

/* RAII marker: while alive, mcx_error() throws instead of exiting. */
struct EntryGuard {
    EntryGuard();
    ~EntryGuard();
};

void record_error(int id, const char* msg);
/* Called by entry points on a caught Error: standalone -> print + exit. */
void handle_error(const Error& e, const Config* cfg);

/* glibc rand()/srand() restated (TYPE_3 additive feedback, published
 * algorithm) so that the per-thread seeds do not depend on the host libc's
 * global state (src/mcx_core.cu:3247-3251, :3381-3387). */
class GlibcRand {
  public:
    explicit GlibcRand(unsigned int seed);
    int next();
    /* n per-thread seed words, each (rand() << 16) | (rand() << 1) |
     * (rand() >> 14) evaluated left to right (src/mcx_core.cu:3381-3387),
     * leaving the generator 3n outputs further on.  Large requests are split
     * into chunks generated on host threads, each chunk's start state
     * reached by jumping ahead (the recurrence is linear over Z/2^32). */
    void seed_words(uint32_t* out, size_t n);
    /* the 31 most recent values, oldest first */
    void state(uint32_t x[31]) const;
    void set_state(const uint32_t x[31]);
  private:
    unsigned int ring_[31];
    int pos_;
};

/* bytes of one detected-photon record (src/mcx_core.cu:2820-2830) */
unsigned int det_reclen(const Config* cfg);
unsigned int det_partialdata(const Config* cfg);

/* rasterize a JSON Shapes description onto cfg->vol (mcx_io.cpp) */
void apply_shapes(Config* cfg, const char* text);

/* host-side encoders of the property-word media formats (mcx_config.cpp) */
unsigned int encode_half2(float a, float b);
unsigned int encode_mua_float(float mua, float unitinmm, bool signed_zero_is_void);

}  // namespace mcx

#define MCX_THROW(id, msg) throw mcx::Error((id), (msg))

#endif
