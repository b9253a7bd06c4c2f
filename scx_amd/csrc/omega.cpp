// Ω generation for the GPU PCA: the reference CPU stream
// ChaCha8Rng::seed_from_u64(seed) sampled with StandardNormal, element i of the
// row-major (rows x cols) matrix = draw i (scx-accel/src/pca/cpu.rs:1174-1178).
//
// Crates restated (absent from the reference tree): rand_chacha 0.3.1
// (ChaCha8, four 16-word blocks per refill, 64-bit counter, zero nonce),
// rand_core 0.6.4 (seed_from_u64 via PCG32, BlockRng next_u32/next_u64),
// rand 0.8.5 (Standard / Open01 f64 conversions), rand_distr 0.4.3
// (StandardNormal ziggurat, 256 boxes, tables printed with 18 decimals).
//
// The ziggurat's rejection steps make the position of draw i depend on every
// earlier draw. The PCA therefore computes, on the device, the keystream and
// the outcome of a draw starting at every position, and the host only walks
// the chain of start positions (plus the rare draws whose value or decision
// depends on libm's exp/log, recomputed here with the host libm); the whole
// matrix is also available here (gaussian_matrix) for the C ABI.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <utility>
#include <vector>

namespace scxpca {

namespace {

struct ChaCha8Stream {
    uint32_t key[8];
    uint64_t block = 0;
    uint32_t words[64];
    unsigned pos = 64;

    explicit ChaCha8Stream(uint64_t seed) {
        // PCG32 expansion of the u64 seed into the 32-byte key.
        uint64_t st = seed;
        for (uint32_t &w : key) {
            st = st * 6364136223846793005ull + 11634580027462260723ull;
            const uint32_t xs = static_cast<uint32_t>(((st >> 18) ^ st) >> 27);
            const uint32_t rot = static_cast<uint32_t>(st >> 59);
            w = (xs >> rot) | (xs << ((32u - rot) & 31u));
        }
    }

    static uint32_t rotl(uint32_t v, int r) { return (v << r) | (v >> (32 - r)); }

    static void quarter(uint32_t *s, int a, int b, int c, int d) {
        s[a] += s[b]; s[d] = rotl(s[d] ^ s[a], 16);
        s[c] += s[d]; s[b] = rotl(s[b] ^ s[c], 12);
        s[a] += s[b]; s[d] = rotl(s[d] ^ s[a], 8);
        s[c] += s[d]; s[b] = rotl(s[b] ^ s[c], 7);
    }

    void refill() {
        for (int blk = 0; blk < 4; ++blk) {
            const uint64_t ctr = block + static_cast<uint64_t>(blk);
            uint32_t init[16] = {0x61707865u, 0x3320646eu, 0x79622d32u, 0x6b206574u,
                                 key[0], key[1], key[2], key[3], key[4], key[5], key[6], key[7],
                                 static_cast<uint32_t>(ctr), static_cast<uint32_t>(ctr >> 32), 0u, 0u};
            uint32_t s[16];
            std::memcpy(s, init, sizeof s);
            for (int dr = 0; dr < 4; ++dr) {
                quarter(s, 0, 4, 8, 12); quarter(s, 1, 5, 9, 13);
                quarter(s, 2, 6, 10, 14); quarter(s, 3, 7, 11, 15);
                quarter(s, 0, 5, 10, 15); quarter(s, 1, 6, 11, 12);
                quarter(s, 2, 7, 8, 13); quarter(s, 3, 4, 9, 14);
            }
            for (int i = 0; i < 16; ++i) words[16 * blk + i] = s[i] + init[i];
        }
        block += 4;
    }

    uint64_t next_u64() {
        if (pos < 63) {
            const uint64_t v = (static_cast<uint64_t>(words[pos + 1]) << 32) | words[pos];
            pos += 2;
            return v;
        }
        if (pos >= 64) {
            refill();
            pos = 2;
            return (static_cast<uint64_t>(words[1]) << 32) | words[0];
        }
        const uint64_t lo = words[63];
        refill();
        pos = 1;
        return (static_cast<uint64_t>(words[0]) << 32) | lo;
    }
};

constexpr int kBoxes = 256;
constexpr double kR = 3.6541528853610088;
constexpr double kV = 0.00492867323399;

struct ZigTables {
    double x[kBoxes + 1];
    double f[kBoxes + 1];

    static double pdf(double v) { return std::exp(-v * v / 2.0); }

    static double decimal18(double v) {
        char text[64];
        std::snprintf(text, sizeof text, "%.18f", v);
        return std::strtod(text, nullptr);
    }

    ZigTables() {
        double raw[kBoxes + 1] = {};
        raw[0] = kV / pdf(kR);
        raw[1] = kR;
        for (int i = 2; i < kBoxes; ++i)
            raw[i] = std::sqrt(-2.0 * std::log(kV / raw[i - 1] + pdf(raw[i - 1])));
        for (int i = 0; i <= kBoxes; ++i) {
            x[i] = decimal18(raw[i]);
            f[i] = decimal18(pdf(raw[i]));
        }
    }
};

const ZigTables &tables() {
    static const ZigTables t;
    return t;
}

inline double bits_to_double(uint64_t b) {
    double d;
    std::memcpy(&d, &b, sizeof d);
    return d;
}

double standard_normal(ChaCha8Stream &rng, const ZigTables &t) {
    constexpr uint64_t kExp2 = static_cast<uint64_t>(1023 + 1) << 52;  // [2,4)
    constexpr uint64_t kExp1 = static_cast<uint64_t>(1023) << 52;      // [1,2)
    for (;;) {
        const uint64_t bits = rng.next_u64();
        const unsigned i = static_cast<unsigned>(bits & 0xffu);
        const double u = bits_to_double((bits >> 12) | kExp2) - 3.0;
        const double xv = u * t.x[i];
        if (std::fabs(xv) < t.x[i + 1]) return xv;
        if (i == 0) {
            double tx = 1.0, ty = 0.0;
            while (-2.0 * ty < tx * tx) {
                const double a = bits_to_double((rng.next_u64() >> 12) | kExp1) - (1.0 - 1.1102230246251565e-16);
                const double b = bits_to_double((rng.next_u64() >> 12) | kExp1) - (1.0 - 1.1102230246251565e-16);
                tx = std::log(a) / kR;
                ty = std::log(b);
            }
            return u < 0.0 ? tx - kR : kR - tx;
        }
        const double unif = static_cast<double>(rng.next_u64() >> 11) * 0x1p-53;
        if (t.f[i + 1] + (t.f[i] - t.f[i + 1]) * unif < ZigTables::pdf(xv)) return xv;
    }
}

}  // namespace

// ---- pieces for the device-side generation ----
// The u64 draws of a stream that only ever calls next_u64 are the pairs
// (u32[2w], u32[2w+1]) of the ChaCha8 keystream (refills keep the read
// position even), so draw w can be computed from its block alone.

void omega_key(uint64_t seed, uint32_t key[8]) {
    const ChaCha8Stream s(seed);
    for (int i = 0; i < 8; ++i) key[i] = s.key[i];
}

void omega_tables(double *x, double *f) {
    const ZigTables &t = tables();
    for (int i = 0; i <= kBoxes; ++i) {
        x[i] = t.x[i];
        f[i] = t.f[i];
    }
}

// One StandardNormal draw starting at u64 draw w of the seed's stream, with
// the host libm: returns the value and the number of u64 draws it consumed.
double omega_draw_at(uint64_t seed, uint64_t w, uint32_t *consumed) {
    ChaCha8Stream rng(seed);
    rng.block = (2 * w / 64) * 4;
    rng.refill();
    rng.pos = static_cast<unsigned>((2 * w) % 64);
    const uint64_t before = rng.block * 16 - (64 - rng.pos);  // u32 words consumed so far
    const double v = standard_normal(rng, tables());
    const uint64_t after = rng.block * 16 - (64 - rng.pos);
    *consumed = static_cast<uint32_t>((after - before) / 2);
    return v;
}

// Walk the chain of draw starts: s_0 = 0, s_{j+1} = s_j + c(s_j), with c from
// the device (bits 0-6: draws consumed; bit 7: recompute on the host, then
// the host value overrides the device one). Returns false when the chain
// leaves the n_words computed on the device.
bool omega_walk(const uint8_t *c, uint64_t n_words, uint64_t n_samples, uint64_t seed, uint32_t *start,
                std::vector<std::pair<uint32_t, double>> *overrides) {
    constexpr uint64_t kOnes = 0x0101010101010101ull;  // eight one-word draws in a row
    uint64_t w = 0;
    for (uint64_t j = 0; j < n_samples; ++j) {
        // fast path: most draws take one word, so eight starts at once when the
        // next eight outcome bytes all say so (the chain is the same either way)
        while (j + 8 <= n_samples && w + 8 <= n_words) {
            uint64_t x;
            std::memcpy(&x, c + w, 8);
            if (x != kOnes) break;
            for (uint32_t t = 0; t < 8; ++t) start[j + t] = static_cast<uint32_t>(w + t);
            j += 8;
            w += 8;
        }
        if (j >= n_samples) break;
        if (w >= n_words) return false;
        start[j] = static_cast<uint32_t>(w);
        const uint8_t cw = c[w];
        if (cw & 0x80u) {
            uint32_t used = 0;
            const double v = omega_draw_at(seed, w, &used);
            overrides->emplace_back(static_cast<uint32_t>(w), v);
            w += used;
        } else {
            w += cw;
        }
    }
    return true;
}

void gaussian_matrix(uint64_t rows, uint64_t cols, uint64_t seed, double *out) {
    const ZigTables &t = tables();
    ChaCha8Stream rng(seed);
    const uint64_t n = rows * cols;
    for (uint64_t i = 0; i < n; ++i) out[i] = standard_normal(rng, t);
}

}  // namespace scxpca
