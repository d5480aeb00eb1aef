// ppf_dispatch.hip -- run-time choice of the line length: the pass
// dispatchers of ppf_plan.hpp call the per-length launchers the size-group
// translation units instantiate.
#include "ppf_plan.hpp"

namespace ppf {

template <typename T>
int dispatch_xfwd(const ppf_plan *p, const double *in, T *W, const AxisTables<T> &tb, bool res, hipStream_t s,
                  int nplanes, ZRows zr) {
    if (nplanes < 0) nplanes = p->d.n3m;
    switch (p->d.n1m) {
#define PPF_CASE(n) \
    case n: return xfwd_n<T, n>(p, in, W, tb, res, zr, s, nplanes);
        PPF_FOR_SIZES(PPF_CASE)
#undef PPF_CASE
    }
    return PPF_ERR_SIZE;
}

template <typename T>
int dispatch_xinv(const ppf_plan *p, const T *W, double *P, const AxisTables<T> &tb, const double *dmx1, bool acc,
                  hipStream_t s, int nplanes) {
    if (nplanes < 0) nplanes = p->Lz;
    switch (p->d.n1m) {
#define PPF_CASE(n) \
    case n: return xinv_n<T, n>(p, W, P, tb, dmx1, acc, s, nplanes);
        PPF_FOR_SIZES(PPF_CASE)
#undef PPF_CASE
    }
    return PPF_ERR_SIZE;
}

template <typename T>
static int dispatch_y_any(const ppf_plan *p, T *W, const AxisTables<T> &tb, bool inverse, const double *dmx2,
                          hipStream_t s, int nplanes) {
    if (nplanes < 0) nplanes = inverse ? p->Lz : p->d.n3m;
    switch (p->d.n2m) {
#define PPF_CASE(n) \
    case n: return y_n<T, n>(p, W, tb, inverse, dmx2, s, nplanes);
        PPF_FOR_SIZES(PPF_CASE)
#undef PPF_CASE
    }
    return PPF_ERR_SIZE;
}

template <typename T>
int dispatch_yfwd(const ppf_plan *p, T *W, const AxisTables<T> &tb, hipStream_t s, int nplanes) {
    return dispatch_y_any<T>(p, W, tb, false, nullptr, s, nplanes);
}
template <typename T>
int dispatch_yinv(const ppf_plan *p, T *W, const AxisTables<T> &tb, const double *dmx2, hipStream_t s, int nplanes) {
    return dispatch_y_any<T>(p, W, tb, true, dmx2, s, nplanes);
}

#define PPF_DISPATCH_INST(T)                                                                                       \
    template int dispatch_xfwd<T>(const ppf_plan *, const double *, T *, const AxisTables<T> &, bool, hipStream_t, \
                                  int, ZRows);                                                                     \
    template int dispatch_xinv<T>(const ppf_plan *, const T *, double *, const AxisTables<T> &, const double *,    \
                                  bool, hipStream_t, int);                                                         \
    template int dispatch_yfwd<T>(const ppf_plan *, T *, const AxisTables<T> &, hipStream_t, int);                \
    template int dispatch_yinv<T>(const ppf_plan *, T *, const AxisTables<T> &, const double *, hipStream_t, int);
PPF_DISPATCH_INST(double)
PPF_DISPATCH_INST(float)
#undef PPF_DISPATCH_INST

}  // namespace ppf
