// Writes std_kats.json: known-answer vectors of the C++ standard library's
// random-number facilities as GNU libstdc++ implements them (the library the
// reference is built with).  Uses the standard library only -- no code from
// the reference.  Build and run with make_std_kats.sh.
#include <algorithm>
#include <cstdint>
#include <cstdio>
#include <numeric>
#include <cstring>
#include <random>
#include <vector>

static void print_seq(const char* name, const std::vector<long long>& v, bool last = false) {
  std::printf("  \"%s\": [", name);
  for (size_t i = 0; i < v.size(); i++) std::printf("%s%lld", i ? ", " : "", v[i]);
  std::printf("]%s\n", last ? "" : ",");
}

int main() {
  const long long kSeed = 27491095;
  std::printf("{\n");
  {  // mt19937 default-seeded 10000th output (standard-mandated value)
    std::mt19937 g;
    g.discard(9999);
    std::printf("  \"mt19937_10000th\": %lu,\n", (unsigned long)g());
  }
  {  // uniform_int_distribution<int32_t>(0, 2^s - 1) from mt19937(kSeed)
    for (int s : {10, 15, 22}) {
      std::mt19937 g(kSeed);
      std::uniform_int_distribution<int32_t> d(0, (1 << s) - 1);
      std::vector<long long> v;
      for (int i = 0; i < 64; i++) v.push_back(d(g));
      char name[64];
      std::snprintf(name, sizeof(name), "uniform_int_pow2_%d", s);
      print_seq(name, v);
    }
    std::mt19937 g(kSeed);
    std::uniform_int_distribution<int32_t> d(0, 131067);  // non power of two
    std::vector<long long> v;
    for (int i = 0; i < 256; i++) v.push_back(d(g));
    print_seq("uniform_int_131068", v);
  }
  {  // uniform_real_distribution<float>(0, 1): bit patterns
    std::mt19937 g(kSeed + 3);
    std::uniform_real_distribution<float> d(0, 1.0f);
    std::vector<long long> v;
    for (int i = 0; i < 256; i++) {
      float f = d(g);
      uint32_t b;
      std::memcpy(&b, &f, 4);
      v.push_back(b);
    }
    print_seq("uniform_float_bits_seed_plus3", v);
  }
  {  // shuffle of the identity, both libstdc++ code paths (n < 2^16 pairs, n >= 2^16 single)
    for (int n : {1, 2, 7, 1000, 1024, 65535, 65536, 70000}) {
      std::vector<int32_t> p(n);
      std::iota(p.begin(), p.end(), 0);
      std::mt19937 g(kSeed);
      std::shuffle(p.begin(), p.end(), g);
      std::vector<long long> v;
      // first 64 and last 64 entries, and a checksum of all
      long long chk = 0;
      for (int i = 0; i < n; i++) chk = (chk * 1000003 + p[i]) % 2147483647;
      for (int i = 0; i < std::min(n, 64); i++) v.push_back(p[i]);
      for (int i = std::max(64, n - 64); i < n; i++) v.push_back(p[i]);
      v.push_back(chk);
      char name[64];
      std::snprintf(name, sizeof(name), "shuffle_%d", n);
      print_seq(name, v);
    }
  }
  {  // GCC evaluation order of f(a(), b()) for the pair constructor idiom
    struct P {
      int a, b;
      P(int x, int y) : a(x), b(y) {}
    };
    int c = 0;
    auto f = [&c] { return c++; };
    P p(f(), f());
    std::printf("  \"pair_ctor_first_arg_value\": %d\n", p.a);
  }
  std::printf("}\n");
  return 0;
}
