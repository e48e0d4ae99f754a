// Forced-include prelude for the stable-order build of the reference harness.
// Pulls in the standard headers first, then maps the identifier `sort` to a
// wrapper around std::stable_sort, so that the reference's per-layer bin
// ordering (src/CA/PrimaryVertexContext.cxx:52-54) becomes stable without
// touching its source.  The wrapper adapts the reference comparator, which
// takes non-const references, to the const arguments std::stable_sort passes.
// TEST INFRASTRUCTURE ONLY.
#include <algorithm>
#include <array>
#include <cmath>
#include <ctime>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <iterator>
#include <memory>
#include <vector>

namespace std {
template <class It, class Cmp>
void itsca_stable_order(It first, It last, Cmp cmp)
{
  typedef typename std::iterator_traits<It>::value_type T;
  std::stable_sort(first, last, [&cmp](const T& a, const T& b) { return cmp(const_cast<T&>(a), const_cast<T&>(b)); });
}
}  // namespace std

#define sort itsca_stable_order
