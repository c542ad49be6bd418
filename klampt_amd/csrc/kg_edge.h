// Sample order of straight-line edge checks, shared by the host code and the
// HIP kernels.
//
// A segment discretised into n steps has interior samples i = 1..n-1 at
// u = i/n (Python/klampt/plan/robotcspace.py:216-228).  They are checked in
// level-order bisection: the intervals [lo, hi) of the current level, left
// to right, each contributes its midpoint lo + (hi-lo)/2 when hi - lo >= 2
// and is split there for the next level, starting from [0, n).
//
// kg_bisect_index(n, k) is the k-th sample of that order (0 <= k < n-1)
// without building the order: at depth d every interval has length
// f = n >> d or f + 1, and (n - f 2^d) of the 2^d intervals have the larger
// length.  A level contributes every interval when f >= 2, only the
// length-2 ones when f == 1, none below.  Within a subtree of length L at
// remaining depth r the intervals of the larger length number L - f 2^r,
// which steers the descent to the j-th contributing interval of the level.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

__host__ __device__ inline int kg_bisect_index(int n, int k) {
  int d = 0;
  int64_t start = 0;
  int f = n;
  for (;;) {
    f = n >> d;
    int64_t cnt;
    if (f >= 2)
      cnt = (int64_t)1 << d;
    else if (f == 1)
      cnt = (int64_t)n - ((int64_t)1 << d);
    else
      return -1;  // k out of range
    if (k < start + cnt) break;
    start += cnt;
    d++;
  }
  int64_t j = k - start;
  int lo = 0, L = n;
  for (int r = d; r > 0; r--) {
    const int Ll = L >> 1;
    bool right;
    if (f >= 2) {
      right = ((j >> (r - 1)) & 1) != 0;
    } else {
      const int64_t big_left = (int64_t)Ll - ((int64_t)f << (r - 1));
      right = j >= big_left;
      if (right) j -= big_left;
    }
    if (right) {
      lo += Ll;
      L -= Ll;
    } else {
      L = Ll;
    }
  }
  return lo + (L >> 1);
}
