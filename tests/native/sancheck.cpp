// sancheck.cpp -- the block codec's own source (zfp-rs_amd/csrc, through
// hostcheck.cpp) and the oracle, built together under AddressSanitizer and
// UndefinedBehaviorSanitizer, run over seeded fields and random streams.
// Test infrastructure only: it checks the kernels' index arithmetic, LDS
// phases (run thread by thread on the CPU) and the oracle for out-of-bounds
// accesses, overflows and invalid shifts, and that both agree bit for bit.
//
// Exit status 0 when every case agrees and no sanitizer reported.
#include "hostcheck.cpp"

#include <math.h>
#include <stdio.h>
#include <string.h>

#include "zfp_oracle.h"

namespace {

uint64_t rng_state = 0x9e3779b97f4a7c15ull;
uint64_t next_u64()
{
  // splitmix64
  uint64_t z = (rng_state += 0x9e3779b97f4a7c15ull);
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
  return z ^ (z >> 31);
}
double next_unit() { return (double)(next_u64() >> 11) * 0x1p-53; }

int failures = 0;

size_t esize(int type) { return type == ORC_INT32 || type == ORC_FLOAT ? 4 : 8; }

// a smooth field of `n` values with some noise; `special` puts infinities,
// NaNs and subnormals in a few places (floats only)
void fill(int type, void *data, size_t n, bool special)
{
  for (size_t i = 0; i < n; i++) {
    const double v = 1000.0 * sin(0.05 * (double)i) + (next_unit() - 0.5);
    switch (type) {
    case ORC_INT32: ((int32_t *)data)[i] = (int32_t)(v * 1000.0); break;
    case ORC_INT64: ((int64_t *)data)[i] = (int64_t)(v * 1e9); break;
    case ORC_FLOAT: ((float *)data)[i] = (float)v; break;
    default: ((double *)data)[i] = v; break;
    }
  }
  if (special && (type == ORC_FLOAT || type == ORC_DOUBLE)) {
    for (size_t i = 0; i < n; i += 97) {
      const double s = i % 3 == 0 ? INFINITY : i % 3 == 1 ? NAN : 1e-310;
      if (type == ORC_FLOAT) ((float *)data)[i] = i % 3 == 2 ? 1e-40f : (float)s;
      else ((double *)data)[i] = s;
    }
  }
}

struct Case {
  int type, dims;
  size_t n[3];
  orc_config cfg;
};

bool is_fixed(const orc_config &c, int type, int dims)
{
  return c.minbits == c.maxbits && orc_block_bits(&c, type, dims) == c.maxbits;
}

void run(const Case &k, bool special, uint32_t start)
{
  const int dims = k.dims;
  size_t count = 1, nb = 1;
  for (int a = 0; a < dims; a++) {
    count *= k.n[a];
    nb *= (k.n[a] + 3) / 4;
  }
  const size_t esz = esize(k.type);
  std::vector<uint8_t> data(count * esz), ref(count * esz), got(count * esz);
  fill(k.type, data.data(), count, special);
  const size_t dz[3] = {k.n[0], dims > 1 ? k.n[1] : 0, dims > 2 ? k.n[2] : 0};
  const size_t cap = orc_maximum_size(&k.cfg, k.type, dz) / 8 + 4 + start / 64 + 1;
  std::vector<uint64_t> ow(cap, 0), hw(cap, 0);
  orc_stream s;
  orc_stream_init(&s, ow.data(), cap);
  for (uint32_t i = 0; i < start; i++) orc_write_bit(&s, i & 1);
  orc_field f{k.type, dz[0], dz[1], dz[2], 0, 0, 0, data.data()};
  int err = 0;
  const size_t bytes = orc_compress(&s, &k.cfg, &f, &err);
  if (err) {
    printf("FAIL oracle compress err %d (type %d dims %d)\n", err, k.type, dims);
    failures++;
    return;
  }
  // the same prefix bits, then the kernels' code
  for (uint32_t i = 0; i < start; i++) hw[i / 64] |= (uint64_t)(i & 1) << (i % 64);
  const int64_t n4[4] = {(int64_t)k.n[0], dims > 1 ? (int64_t)k.n[1] : 1, dims > 2 ? (int64_t)k.n[2] : 1, 1};
  const int64_t s4[4] = {1, n4[0], n4[0] * n4[1], n4[0] * n4[1] * n4[2]};
  const HcCfg hc{k.cfg.minbits, k.cfg.maxbits, k.cfg.maxprec, k.cfg.minexp, k.cfg.rounding, k.cfg.tight};
  const int rev = orc_is_reversible(&k.cfg);
  const uint32_t fixed = is_fixed(k.cfg, k.type, dims) ? k.cfg.maxbits : 0;
  std::vector<uint32_t> lens(nb);
  const uint32_t variants[3] = {0, 3, 64};
  for (uint32_t T : variants) {
    if (fixed > 4096 && T) continue;
    std::fill(hw.begin() + start / 64 + (start % 64 ? 1 : 0), hw.end(), 0ull);
    if (start % 64) hw[start / 64] &= (1ull << (start % 64)) - 1;
    const uint64_t end = hc_compress(k.type, dims, rev, data.data(), n4, s4, &hc, hw.data(), cap, start,
                                     fixed, lens.data(), T, 0);
    if ((end + 63) / 64 * 8 != bytes || memcmp(hw.data(), ow.data(), bytes) != 0) {
      printf("FAIL stream (type %d dims %d T %u fixed %u start %u special %d)\n", k.type, dims, T,
             fixed, start, special);
      failures++;
      return;
    }
    // decode: the oracle, then the kernels' code
    orc_stream r;
    orc_stream_init(&r, ow.data(), cap);
    orc_rseek(&r, start);
    orc_field g{k.type, dz[0], dz[1], dz[2], 0, 0, 0, ref.data()};
    orc_decompress(&r, &k.cfg, &g, &err);
    hc_decompress(k.type, dims, rev, got.data(), n4, s4, &hc, hw.data(), cap, start, fixed,
                  fixed ? nullptr : lens.data(), T);
    if (memcmp(ref.data(), got.data(), count * esz) != 0) {
      printf("FAIL values (type %d dims %d T %u fixed %u start %u special %d)\n", k.type, dims, T,
             fixed, start, special);
      failures++;
      return;
    }
  }
}

// a random stream decoded by both (fixed-rate configs: every block's bits)
void run_random_stream(int type, int dims, const orc_config &c)
{
  const size_t n[3] = {13, dims > 1 ? (size_t)9 : 0, dims > 2 ? (size_t)7 : 0};
  size_t count = 1, nb = 1;
  for (int a = 0; a < dims; a++) {
    count *= n[a];
    nb *= (n[a] + 3) / 4;
  }
  const size_t nwords = nb * (c.maxbits + 63) / 64 + 4;
  std::vector<uint64_t> w(nwords);
  for (auto &x : w) x = next_u64() & next_u64() & next_u64();  // sparse bits: long runs
  const size_t esz = esize(type);
  std::vector<uint8_t> ref(count * esz), got(count * esz);
  orc_stream r;
  orc_stream_init(&r, w.data(), nwords);
  orc_field g{type, n[0], n[1], n[2], 0, 0, 0, ref.data()};
  int err = 0;
  orc_decompress(&r, &c, &g, &err);
  const int64_t n4[4] = {(int64_t)n[0], dims > 1 ? (int64_t)n[1] : 1, dims > 2 ? (int64_t)n[2] : 1, 1};
  const int64_t s4[4] = {1, n4[0], n4[0] * n4[1], n4[0] * n4[1] * n4[2]};
  const HcCfg hc{c.minbits, c.maxbits, c.maxprec, c.minexp, c.rounding, c.tight};
  hc_decompress(type, dims, orc_is_reversible(&c), got.data(), n4, s4, &hc, w.data(), nwords, 0,
                c.maxbits, nullptr, 0);
  if (memcmp(ref.data(), got.data(), count * esz) != 0) {
    printf("FAIL random stream (type %d dims %d maxbits %u)\n", type, dims, c.maxbits);
    failures++;
  }
}

}  // namespace

int main()
{
  const int types[4] = {ORC_INT32, ORC_INT64, ORC_FLOAT, ORC_DOUBLE};
  const size_t shapes[3][3] = {{37, 0, 0}, {13, 9, 0}, {9, 6, 7}};
  int cases = 0;
  for (int ti = 0; ti < 4; ti++) {
    const int type = types[ti];
    for (int dims = 1; dims <= 3; dims++) {
      std::vector<orc_config> cfgs;
      for (double rate : {2.0, 8.0, 13.0, 32.0}) {
        const uint32_t bits = orc_rate_bits(rate, type, dims, 0);
        cfgs.push_back(orc_config{bits, bits, 64, -1074, ORC_ROUND_NEVER, 0});
      }
      cfgs.push_back(orc_config{1, 16658, 16, -1074, ORC_ROUND_FIRST, 0});
      cfgs.push_back(orc_config{1, 16658, 64, -20, ORC_ROUND_LAST, 1});
      cfgs.push_back(orc_config{1, 16658, 64, -1075, ORC_ROUND_NEVER, 0});
      cfgs.push_back(orc_config{100, 600, 20, -30, ORC_ROUND_NEVER, 0});
      for (const orc_config &c : cfgs) {
        Case k{type, dims, {shapes[dims - 1][0], shapes[dims - 1][1], shapes[dims - 1][2]}, c};
        for (uint32_t start : {0u, 13u}) {
          run(k, false, start);
          run(k, true, start);
          cases += 2;
        }
        if (is_fixed(c, type, dims)) {
          for (int r = 0; r < 4; r++) run_random_stream(type, dims, c);
          cases += 4;
        }
      }
    }
  }
  if (failures) {
    printf("%d of %d cases failed\n", failures, cases);
    return 1;
  }
  printf("sancheck: %d cases agree\n", cases);
  return 0;
}
