// kahan.hpp -- compensated summation shared by the PageRank kernels.
#pragma once
#include <hip/hip_runtime.h>

namespace byo {

// Running sum with its Kahan compensation: value = s - c.
template <class T>
struct Acc {
  T s, c;
};

// Kahan step (per lane).
template <class T>
__device__ __forceinline__ void kahan_add(Acc<T> &a, T x) {
  T y = x - a.c;
  T t = a.s + y;
  a.c = (t - a.s) - y;
  a.s = t;
}

// Compensated combine of two partial sums (TwoSum on the heads).
template <class T>
__device__ __forceinline__ Acc<T> combine(Acc<T> a, Acc<T> b) {
  T s = a.s + b.s;
  T bp = s - a.s;
  T err = (a.s - (s - bp)) + (b.s - bp);
  // Kahan carries the negated error in c (value = s - c)
  return Acc<T>{s, a.c + b.c - err};
}

}  // namespace byo
