// sap.h -- fused SAP half-sweep kernel (one workgroup per block of one colour).
#pragma once

#include "kernels.h"

namespace qws {


enum SapFlags {
  SAP_EPI_SUPDATE = 0,   // s_c = s_c + in_c - x + kappa T^-1 H_in x ; q_c = x
  SAP_EPI_OUT = 1,       // out_c = x - kappa T^-1 H_in x            ; q_c = x
  SAP_EPI_OUT_FACE = 2,  // out_c = -kappa T^-1 H_face q + x - kappa T^-1 H_in x ; q_c = x
  SAP_EPI_NONE = 3,      // q_c = x only
  SAP_EPI_MASK = 3,
  SAP_FACE_PRO = 4,      // prologue: s_c += kappa T^-1 H_face qface
  SAP_SRC_IN = 8,        // take s_c from `in` (first sweep: s = in)
  SAP_FACE_OUT = 16,     // also write the face buffer (FP32 slots) for the next launch's prologue
  SAP_FACE_FB = 32,      // the prologue reads the face buffer instead of gathering qface
  SAP_SRC_PUPD = 64,     // with SAP_SRC_IN (FP32): first in_c = r_c + beta (in_c - omega q_c), stored
};

// The BiCGStab direction update folded into the first sweep (SAP_SRC_PUPD): `in` is
// the direction p itself, updated in place from r, q and the RHS's omega, beta.
struct SapSrcUpdate {
  float* p = nullptr;
  const float* r = nullptr;
  const float* q = nullptr;
  const Scalars* sc = nullptr;
};

// One lattice's share of a launch fused over several lattices: what a single launch
// takes as arguments, per lattice.  For the face-term kernel, kappa holds its factor.
template <typename S> struct SapPart {
  Geo g;
  const int* active;
  long long stride;
  int nrhs;
  float kappa;
  S* s;
  const S* in;
  const S* qface;
  S* qout;
  S* out;
  const S* U;
  const S* C;
  float* fb;
  long long fbs;
  SapSrcUpdate upd;
};
// A launch fused over several lattices: workgroup w runs block table[w].y of part
// table[w].x (device arrays).  The parts share the block volume, the link storage,
// the flags and nm; xcd: map the workgroups to the table XCD by XCD.
template <typename S> struct SapMulti {
  const SapPart<S>* parts = nullptr;
  const int2* table = nullptr;
  int n = 0;
  int xcd = 1;
};

bool sap_fusable(const Geo& g);
// floats per RHS of the face buffer, or 0 when the block shape does not allow it
// (its slots must fit the iterate's LDS: 2 sum_mu Vb / B_mu <= 2 Vb)
long long sap_face_buffer_floats(const Geo& g);
// out += gamma T^-1 H_face q on the blocks of `sel` (FP32), H_face q from the face
// buffer that the preceding SAP launch of the other colour wrote (SAP_FACE_OUT)
void launch_sap_face(const Geo& g, BlockSel sel, const Batch& B, int nl, float gamma, float* out, const float* fb,
                     long long fbs, const float* U, const float* C, hipStream_t st);
template <typename S>
void launch_sap_block(const Geo& g, BlockSel sel, const Batch& B, int nl, float kappa, int nm, int flags, S* s,
                      const S* in, const S* qface, S* qout, S* out, const S* U, const S* C, float* fb,
                      long long fbs, hipStream_t st, const SapSrcUpdate& upd = SapSrcUpdate());
// The same over several lattices in one launch (FP32 storage): vb = their sites per
// block, nl their FP32 link storage; every part has a face buffer when `flags` use one.
void launch_sap_multi(const SapMulti<float>& m, int vb, int nl, int nm, int flags, hipStream_t st);
// launch_sap_face over several lattices (part.out, .fb, .fbs, .U, .C; factor in .kappa)
void launch_sap_face_multi(const SapMulti<float>& m, int vb, int nl, hipStream_t st);

}  // namespace qws
