// radix.hpp -- parallel, stable LSD radix sort of (key, payload) pairs.
#pragma once

#include <omp.h>

#include <cstdint>
#include <cstring>
#include <vector>

namespace blco {

// Sorts keys[0..n) ascending over their low `bits` bits, permuting `idx`
// alongside.  Stable: equal keys keep their input order.  K is uint64_t or
// unsigned __int128.
template <typename K, typename P>
void radix_sort(uint64_t n, K* keys, P* idx, int bits)
{
    if (n < 2 || bits <= 0) return;
    constexpr int DIG = 11, NB = 1 << DIG;
    std::vector<K> k2(n);
    std::vector<P> i2(n);
    int nt = omp_get_max_threads();
    if ((uint64_t)nt > n / 4096 + 1) nt = (int)(n / 4096 + 1);
    std::vector<uint64_t> hist((size_t)nt * NB);
    K* sk = keys; P* si = idx;
    K* dk = k2.data(); P* di = i2.data();
    for (int sh = 0; sh < bits; sh += DIG) {
        std::memset(hist.data(), 0, hist.size() * sizeof(uint64_t));
        #pragma omp parallel num_threads(nt)
        {
            int t = omp_get_thread_num();
            uint64_t lo = n * (uint64_t)t / nt, hi = n * (uint64_t)(t + 1) / nt;
            uint64_t* h = hist.data() + (size_t)t * NB;
            for (uint64_t i = lo; i < hi; ++i) h[(unsigned)(sk[i] >> sh) & (NB - 1)]++;
        }
        uint64_t run = 0;
        for (int d = 0; d < NB; ++d)
            for (int t = 0; t < nt; ++t) {
                uint64_t c = hist[(size_t)t * NB + d];
                hist[(size_t)t * NB + d] = run;
                run += c;
            }
        #pragma omp parallel num_threads(nt)
        {
            int t = omp_get_thread_num();
            uint64_t lo = n * (uint64_t)t / nt, hi = n * (uint64_t)(t + 1) / nt;
            uint64_t* h = hist.data() + (size_t)t * NB;
            for (uint64_t i = lo; i < hi; ++i) {
                uint64_t o = h[(unsigned)(sk[i] >> sh) & (NB - 1)]++;
                dk[o] = sk[i];
                di[o] = si[i];
            }
        }
        std::swap(sk, dk);
        std::swap(si, di);
    }
    if (sk != keys) {
        std::memcpy(keys, sk, n * sizeof(K));
        std::memcpy(idx, si, n * sizeof(P));
    }
}

}  // namespace blco
