// sap.hip -- fused SAP half-sweep: one workgroup per SAP block of one colour.
//
// One launch does, for every block of colour c (multiplicative Schwarz step of
// qws.cc:809-841, with the face update of the previous half-sweep moved in front):
//   prologue   s_c += kappa T^-1 H_face q            (ddd_out_pre_s_/ddd_out_pos_s_)
//   block solve x = A_c s_c: x_0 = s_c, x_{k+1} = s_c + kappa T^-1 H_in x_k, nm times
//              (jinv_ddd_in_s_, static_solver.cc:65-140)
//   epilogue   s_c += in_c - x + kappa T^-1 H_in x,  q_c = x   (qws.cc:814-823)
//              or the final sequences of prec_ddd_s_ / prec_s_ (qws.cc:842-925, 967-972)
// The block's forward links and the iterate x live in LDS, the site's clover
// inverse in registers, so links and clover cross HBM once per half-sweep instead
// of once per operator application.  Requires proper colouring (face neighbours
// belong to the other colour) and 256 or 512 sites per block (one thread per site).
// Instantiated for FP32 storage and for FP16 storage (FP16 operands, FP32
// accumulation).
#include <type_traits>

#include "kernels.h"
#include "sap.h"

namespace qws {

namespace sap256 {
constexpr int VB = 256;  // 4^4 blocks
#include "sap_kernels.inc"
}  // namespace sap256
namespace sap512 {
constexpr int VB = 512;  // e.g. 8x4x4x4 blocks (FP32: 144 KB of LDS per workgroup)
#include "sap_kernels.inc"
}  // namespace sap512

bool sap_fusable(const Geo& g) { return g.Vb == 256 || g.Vb == 512; }

long long sap_face_buffer_floats(const Geo& g) {
  if (!sap_fusable(g)) return 0;
  const int S = g.Vb == 512 ? sap512::face_slots(g) : sap256::face_slots(g);
  if (S > 2 * g.Vb) return 0;
  return (long long)g.nblk * S * 12;
}

template <typename S>
void launch_sap_block(const Geo& g, BlockSel sel, const Batch& B, int nl, float kappa, int nm, int flags, S* s,
                      const S* in, const S* qface, S* qout, S* out, const S* U, const S* C, float* fb,
                      long long fbs, hipStream_t st, const SapSrcUpdate& upd) {
  if (sel.n == 0) return;
  if (nm < 1) nm = 1;  // the first application is unconditional (static_solver.cc:77-101)
  if ((flags & (SAP_FACE_OUT | SAP_FACE_FB)) && (!fb || sap_face_buffer_floats(g) == 0))
    throw std::invalid_argument("fused SAP: face-buffer flags need a buffer and a block shape that allows one");
  if (flags & SAP_SRC_PUPD) {
    if (!std::is_same<S, float>::value || !(flags & SAP_SRC_IN) || !upd.p || !upd.r || !upd.q || !upd.sc ||
        (const void*)upd.p != (const void*)in)
      throw std::invalid_argument("fused SAP: the source update needs FP32, SAP_SRC_IN and in == p, r, q, scalars");
  }
  if (g.Vb == 512) sap512::launch<S>(g, sel, B, nl, kappa, nm, flags, s, in, qface, qout, out, U, C, fb, fbs, st, upd);
  else sap256::launch<S>(g, sel, B, nl, kappa, nm, flags, s, in, qface, qout, out, U, C, fb, fbs, st, upd);
}

void launch_sap_face(const Geo& g, BlockSel sel, const Batch& B, int nl, float gamma, float* out, const float* fb,
                     long long fbs, const float* U, const float* C, hipStream_t st) {
  if (sel.n == 0) return;
  if (!fb || sap_face_buffer_floats(g) == 0) throw std::invalid_argument("face term: no face buffer for this lattice");
  if (g.Vb == 512) sap512::launch_face(g, sel, B, nl, gamma, out, fb, fbs, U, C, st);
  else sap256::launch_face(g, sel, B, nl, gamma, out, fb, fbs, U, C, st);
}

void launch_sap_multi(const SapMulti<float>& m, int vb, int nl, int nm, int flags, hipStream_t st) {
  if (m.n == 0) return;
  if (nm < 1) nm = 1;  // the first application is unconditional (static_solver.cc:77-101)
  if (flags & SAP_SRC_PUPD) {
    if (!(flags & SAP_SRC_IN)) throw std::invalid_argument("fused SAP: the source update needs SAP_SRC_IN");
  }
  if (vb == 512) sap512::launch_multi(m, nl, nm, flags, st);
  else if (vb == 256) sap256::launch_multi(m, nl, nm, flags, st);
  else throw std::invalid_argument("fused SAP: 256 or 512 sites per block");
}

void launch_sap_face_multi(const SapMulti<float>& m, int vb, int nl, hipStream_t st) {
  if (m.n == 0) return;
  if (vb == 512) sap512::launch_face_multi(m, nl, st);
  else if (vb == 256) sap256::launch_face_multi(m, nl, st);
  else throw std::invalid_argument("face term: 256 or 512 sites per block");
}

template void launch_sap_block<float>(const Geo&, BlockSel, const Batch&, int, float, int, int, float*, const float*,
                                      const float*, float*, float*, const float*, const float*, float*, long long,
                                      hipStream_t, const SapSrcUpdate&);
template void launch_sap_block<__half>(const Geo&, BlockSel, const Batch&, int, float, int, int, __half*,
                                       const __half*, const __half*, __half*, __half*, const __half*, const __half*,
                                       float*, long long, hipStream_t, const SapSrcUpdate&);

}  // namespace qws
