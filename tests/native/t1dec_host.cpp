// Host build of the decoder kernel's per-block code (libgail_amd/csrc/t1_decode_block.h) for
// the CPU tests: the same source the GPU runs, with plain arrays in place of LDS and plain
// additions in place of atomic adds, so its words can be checked against the oracle's
// decoder without a GPU. Test infrastructure: loaded by tests/ only.
#include <cstring>
#include <vector>

#include "t1_decode_block.h"
#include "t1_tables.h"

using namespace gail_amd;

namespace {

struct HostState {
    uint32_t m[4][32];
    uint32_t c[19];
    const uint32_t* mq;
    const uint8_t* zct;  // Table D.1 of the block's orientation
    void reset() {
        std::memset(m, 0, sizeof m);
        for (uint32_t i = 0; i < 19; ++i) c[i] = mq[t1d_initial_state(i)];
    }
    uint32_t rows(int t, uint32_t r) const { return m[t][r]; }
    void set_rows(int t, uint32_t r, uint32_t v) { m[t][r] = v; }
    uint32_t ctx(uint32_t i) const { return c[i]; }
    void set_ctx(uint32_t i, uint32_t v) { c[i] = v; }
    uint32_t tab(uint32_t i) const { return mq[i]; }
    uint32_t zc(uint32_t code) const { return zct[code]; }
};

struct HostSink {
    uint32_t* base;
    uint32_t stride;
    uint32_t limit;
    int bad;
    void add(uint32_t i, uint32_t v) {
        if (i >= limit) {
            bad = 1;
            return;
        }
        base[i] += v;
    }
};

struct Tables {
    uint8_t zc[4 * 128];
    uint32_t mq[47];
    Tables() {
        build_zc_table(zc);
        build_mq_table(mq);
    }
};
const Tables g_tab;

}  // namespace

extern "C" {

// Decodes a codeword placed `start` bytes into a buffer padded as the device buffer is, and
// returns the words with row stride w. Returns 1 when the block wrote outside its 32 x 32
// words, else 0.
int host_t1_decode_words(const uint8_t* data, uint32_t len, uint32_t start, int w, int h, int orient, int nbps,
                         int passes, uint32_t* words) {
    std::vector<uint8_t> buf((size_t)start + len + 64, 0xA5);
    if (len) std::memcpy(buf.data() + start, data, len);
    uint32_t out[32 * 32] = {};
    HostSink sink{out, 32u, 32u * 32u, 0};
    HostState st;
    st.mq = g_tab.mq;
    st.zct = g_tab.zc + 128 * (orient & 3);
    t1d_decode_block(st, buf.data(), start, len, (uint32_t)w, (uint32_t)h, (uint32_t)nbps, (uint32_t)passes, sink);
    for (int y = 0; y < h; ++y) std::memcpy(words + (size_t)y * w, out + (size_t)y * 32, sizeof(uint32_t) * w);
    return sink.bad;
}

}  // extern "C"
