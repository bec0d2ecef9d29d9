// Process-wide launch parameters chosen by the autotuner (or defaults):
// integer values under string keys, read by the launchers.
//
//   mul.q         columns per lane group of the classical kernel (4, 8 or 16)
//   mul.rpark     1: park the first column group in registers, not LDS (Q >= 8)
//   mul.mfma      1: int8 matrix-core classical kernel (balanced 8-bit digits)
//   mul.mfma_wpi  waves per instance of that kernel (0: a quarter of the column tiles)
//   mul.mfma_ipb  instances per workgroup of that kernel (0: 4 / waves per instance)
//   mul.wave16    1 (default): with mul.mfma, 64- and 128-limb operands run one product
//                 per wave on the 16 x 16 x 64 instruction (mul_wave16.hip); with
//                 poly.fused, so does the poly composite (one instance per wave)
//   mul.m16       1 (default): with mul.mfma, other sizes run one product per workgroup on
//                 the 16 x 16 x 64 instruction (mul_mfma16.hip); 0: 32 x 32 x 32 tiles
//   mul.m16_waves waves per product of that kernel (0: from the number of 4-tile passes)
//   mul.prefetch  1: double-buffer the operand LDS reads
//   mul.digits    1: 28-bit-digit classical kernel (one mad per product)
//   mul.threads   threads per block of the word kernel (64..256)
//   add.rows      rows of 128 limbs per wave for small-m additions (1..8)
//   add.threads   threads per block for small-m additions (64..1024)
//   add.nt        1: non-temporal (streaming) 16-byte loads and stores
//   poly.fused    1: the poly composite as one matrix-core kernel (0: a chain of
//                 launches through context scratch)
//   ntt.fast      1: register-pass NTT kernel for transforms of 2^10..2^14 points
//   ntt.vals      values per thread of that kernel, per transform length
//                 ("ntt.vals@L"): 16 or 32, 0 for the built-in choice
//   ragged.threads  threads of the persistent kernel for ops of a size class
//                 ("ragged.threads@C"): 256 or 512
//   ragged.q      columns per lane group of that kernel's products, per size
//                 class ("ragged.q@C"): 4 or 8
//   ragged.mfma_waves waves per workgroup of the ragged matrix-core node of a
//                 size class ("ragged.mfma_waves@C"): 1, 2, 4 or 8
//   ragged.mfma_min ragged multiplications whose shorter operand has at least
//                 this many limbs run on the matrix-core kernel (0: none)
//   ragged.overlap  1: the persistent kernel runs beside the matrix-core node
//   ragged.m16    which ragged products run on 16 x 16 x 64 tiles, per size class: 0
//                 (default, fastest on the 50000-op batch) none, 1 all (and the serving
//                 loop's matrix-core node), 2 those whose shorter operand has at least
//                 1 / ragged.m16_ratio (default 2) of the longer one's limbs; the rest
//                 on 32 x 32 x 32 tiles
//   ragged.m16_waves waves per workgroup of the 16 x 16 x 64 nodes, per size class (default 4)
//   ragged.wave   1 (default): the additions, and the multiplications whose shorter
//                 operand has at most 128 limbs, run one op per wave (16 x 16 x 64 tiles
//                 for the products); the other multiplications on the matrix-core nodes,
//                 no persistent kernel. 0: the persistent kernels and ragged.mfma_min
//   ragged.wave_split 1: the one-per-wave products with a shorter operand of at most
//                 70 limbs on a kernel of their own (fewer registers), a second node
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <mutex>
#include <string>
#include <unordered_map>

#include "midint.h"
#include "midint_internal.h"

namespace midint {

namespace {
std::mutex g_mu;
std::unordered_map<std::string, int64_t>& table() {
    static std::unordered_map<std::string, int64_t> t = {
        {"mul.q", 8}, {"mul.prefetch", 0}, {"mul.digits", 0}, {"mul.threads", 256}, {"mul.rpark", 1}, {"mul.mfma", 0}, {"mul.mfma_wpi", 0}, {"mul.mfma_ipb", 0}, {"mul.wave16", 1}, {"mul.m16", 1}, {"mul.m16_waves", 0}, {"add.rows", 2}, {"add.threads", 256}, {"add.nt", 1},
        {"poly.fused", 1}, {"ntt.fast", 1}, {"ntt.vals", 0}, {"ragged.threads", 256}, {"ragged.q", 4}, {"ragged.mfma_waves", 4}, {"ragged.mfma_min", 32}, {"ragged.overlap", 1}, {"ragged.m16", 0}, {"ragged.m16_ratio", 2}, {"ragged.m16_waves", 4}, {"ragged.wave", 1}, {"ragged.wave_split", 0},
    };
    return t;
}
}  // namespace

int64_t tune_get_class(const char* key, uint32_t limbs, int64_t dflt) {
    uint32_t c = 1;
    while (c < limbs) c <<= 1;
    std::string k = std::string(key) + "@" + std::to_string(c);
    {
        std::lock_guard<std::mutex> lk(g_mu);
        auto it = table().find(k);
        if (it != table().end()) return it->second;
    }
    return tune_get(key, dflt);
}

int64_t tune_get(const char* key, int64_t dflt) {
    std::lock_guard<std::mutex> lk(g_mu);
    auto it = table().find(key);
    return it == table().end() ? dflt : it->second;
}

}  // namespace midint

extern "C" {

// "name" sets the default of a parameter; "name@C" (C a power of two) sets
// it for operand sizes in (C/2, C] limbs.
midint_status midint_tune_set(const char* key, int64_t value) {
    if (!key) return MIDINT_EINVAL;
    std::lock_guard<std::mutex> lk(midint::g_mu);
    auto& t = midint::table();
    std::string k(key);
    const size_t at = k.find('@');
    const std::string base = at == std::string::npos ? k : k.substr(0, at);
    if (t.find(base) == t.end()) return MIDINT_ENOTFOUND;
    if (at != std::string::npos) {
        const std::string cls = k.substr(at + 1);
        if (cls.empty() || cls.find_first_not_of("0123456789") != std::string::npos) return MIDINT_EINVAL;
        const unsigned long c = strtoul(cls.c_str(), nullptr, 10);
        if (c == 0 || (c & (c - 1)) != 0) return MIDINT_EINVAL;
    }
    t[k] = value;
    return MIDINT_OK;
}

midint_status midint_tune_load(const char* path) {
    if (!path) return MIDINT_EINVAL;
    FILE* f = fopen(path, "r");
    if (!f) return MIDINT_ENOTFOUND;
    char line[256];
    midint_status st = MIDINT_OK;
    while (fgets(line, sizeof line, f)) {
        char key[128];
        long long v = 0;
        if (line[0] == '#' || line[0] == '\n') continue;
        if (sscanf(line, "%127s %lld", key, &v) != 2) {
            st = MIDINT_EINVAL;
            continue;
        }
        midint_status s = midint_tune_set(key, v);
        if (s != MIDINT_OK) st = s;
    }
    fclose(f);
    return st;
}

midint_status midint_tune_get(const char* key, int64_t* value) {
    if (!key || !value) return MIDINT_EINVAL;
    std::lock_guard<std::mutex> lk(midint::g_mu);
    auto& t = midint::table();
    auto it = t.find(key);
    if (it == t.end()) return MIDINT_ENOTFOUND;
    *value = it->second;
    return MIDINT_OK;
}

}  // extern "C"
