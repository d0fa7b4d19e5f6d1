// sg_dispatch.cpp -- species-count dispatch of the kernel launchers.
//
// Each supported species count is compiled into its own object
// (sg_kernels.hip with -DSG_NS=N); this file selects the instantiation.
#include "sg_internal.h"

namespace sg {

#define SG_NS_LIST(X) \
  X(2) X(3) X(4) X(5) X(6) X(7) X(8) X(9) X(10) X(11) X(12) X(13) X(14) X(15) X(16) X(17) X(18)

template <int NS> void launch_tempin_t(const Params &, const sg_config *, const Fields &, hipStream_t);
template <int NS> void launch_boundt_t(const Launch &, hipStream_t);
template <int NS> void launch_bountt_t(const Launch &, hipStream_t);
template <int NS> void launch_prim_t(const Launch &, hipStream_t);
template <int NS> void launch_pass_t(int, const Launch &, hipStream_t);
template <int NS> void launch_nscbc_t(const Launch &, hipStream_t);
template <int NS> void launch_radn_t(const Launch &, hipStream_t);
template <int NS> void launch_final_t(const Launch &, int, double *, hipStream_t);

bool species_supported(int ns) {
#define CASE(N) if (ns == N) return true;
  SG_NS_LIST(CASE)
#undef CASE
  return false;
}

void launch_tempin(const Params &P, const sg_config *c, const Fields &F, hipStream_t st) {
  switch (P.ns) {
#define CASE(N) case N: launch_tempin_t<N>(P, c, F, st); break;
    SG_NS_LIST(CASE)
#undef CASE
    default: break;
  }
}
void launch_boundt(const Launch &L, hipStream_t st) {
  switch (L.ns) {
#define CASE(N) case N: launch_boundt_t<N>(L, st); break;
    SG_NS_LIST(CASE)
#undef CASE
    default: break;
  }
}
void launch_bountt(const Launch &L, hipStream_t st) {
  switch (L.ns) {
#define CASE(N) case N: launch_bountt_t<N>(L, st); break;
    SG_NS_LIST(CASE)
#undef CASE
    default: break;
  }
}
void launch_prim(const Launch &L, hipStream_t st) {
  switch (L.ns) {
#define CASE(N) case N: launch_prim_t<N>(L, st); break;
    SG_NS_LIST(CASE)
#undef CASE
    default: break;
  }
}
void launch_pass(int dir, const Launch &L, hipStream_t st) {
  switch (L.ns) {
#define CASE(N) case N: launch_pass_t<N>(dir, L, st); break;
    SG_NS_LIST(CASE)
#undef CASE
    default: break;
  }
}
void launch_radn(const Launch &L, hipStream_t st) {
  switch (L.ns) {
#define CASE(N) case N: launch_radn_t<N>(L, st); break;
    SG_NS_LIST(CASE)
#undef CASE
    default: break;
  }
}
void launch_nscbc(const Launch &L, hipStream_t st) {
  switch (L.ns) {
#define CASE(N) case N: launch_nscbc_t<N>(L, st); break;
    SG_NS_LIST(CASE)
#undef CASE
    default: break;
  }
}
void launch_final(const Launch &L, int irk, double *out, hipStream_t st) {
  switch (L.ns) {
#define CASE(N) case N: launch_final_t<N>(L, irk, out, st); break;
    SG_NS_LIST(CASE)
#undef CASE
    default: break;
  }
}

}  // namespace sg
