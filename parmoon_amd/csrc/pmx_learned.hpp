// pmx_learned.hpp -- the learned post-smoother: a 4-layer 3x3x3 convolution
// network (16 -> 16 -> 16 -> 16 -> 1 channels, GELU between layers) that maps
// a diagonally scaled defect to a correction on the node lattice of a level.
// Input channels per node: [u, a(p,p+e)/a(p,p) for e = +x,-x,+y,-y,+z,-z,
// a(p,p+2e)/a(p,p) for the same six directions (0 where the row has no such
// coupling), 0, 0, 0] -- the scaled defect and the couplings of the row along
// the axes, which tell the network the local convection direction and strength.
// BF16 weights and activations, FP32 accumulation on MFMA
// (v_mfma_f32_16x16x32_bf16), hand-written forward, backward and Adam kernels.
// No counterpart in the reference (its smoothers are MGLevel3D.C:280-563).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

namespace pmx {

constexpr int kNetLayers = 4;
constexpr int kNetCh = 16;   // hidden channels
constexpr int kNetIn = 16;   // input channels: scaled defect + 12 operator features + 3 zero
constexpr int kTaps = 27;    // 3x3x3 stencil of a layer

// layer l maps cin(l) -> cout(l) channels
inline int net_cin(int l) { return l == 0 ? kNetIn : kNetCh; }
inline int net_cout(int l) { return l == kNetLayers - 1 ? 1 : kNetCh; }
// FP32 master parameters of layer l: W[cout][tap][cin] then b[cout]
inline int net_wsize(int l) { return net_cout(l) * kTaps * net_cin(l); }
inline int net_psize(int l) { return net_wsize(l) + net_cout(l); }
int net_total_params();
int net_param_offset(int l);

struct LearnedNet {
  bool ready = false;
  float *params = nullptr;   // all layers, FP32 master copy
  float *grads = nullptr;
  float *adam_m = nullptr, *adam_v = nullptr;
  uint16_t *wbf = nullptr;   // per layer: BF16 weights padded to 16 outputs [16][27][cin]
  uint16_t *wtbf = nullptr;  // per layer: transposed, tap-flipped BF16 weights (input gradient)
  int adam_t = 0;
  // activations for one lattice of N nodes (allocated for the largest use)
  int64_t cap = 0;
  uint16_t *act[kNetLayers] = {nullptr};  // layer inputs, BF16, [N][cin]
  float *pre[kNetLayers] = {nullptr};     // pre-activations (training), [N][cout]
  uint16_t *gbuf[2] = {nullptr, nullptr}; // BF16 gradients flowing backwards, [N][16]
  float *part = nullptr;                  // weight-gradient partial sums
  int64_t part_cap = 0;
  float *out = nullptr;                   // network output [N]
  double *lpart = nullptr;                // loss partial sums
  void release();
};

// device buffers for a training set: `ns` samples on an M^3 lattice
struct LearnedData {
  int M = 0, ns = 0;
  float *u = nullptr;   // scaled input  (D^{-1} d / s)
  float *t = nullptr;   // scaled target (e / s)
  void release();
};

void net_init(LearnedNet &net, uint32_t seed, hipStream_t st);
// returns true if buffers were (re)allocated: launch graphs that captured the
// old activation buffers must then be rebuilt
bool net_reserve(LearnedNet &net, int64_t N, bool training);
// refresh BF16 copies from the FP32 master parameters
void net_sync_bf16(LearnedNet &net, hipStream_t st);
// forward on an M^3 lattice: in[N][kNetIn] (FP32) -> net.out[N] (FP32); training keeps pre-activations
void net_forward(LearnedNet &net, int M, const float *in, bool training, hipStream_t st);
// the same from the BF16 input already in net.act[0]
void net_forward_act(LearnedNet &net, int M, bool training, hipStream_t st);
// inference from net.act[0] straight into the correction: c = out * (*nrm) /
// sqrt_count on interior nodes, 0 on the boundary (the output layer applies
// the scaling and the boundary mask to its FP32 result)
// the first layer's input built inside that layer (k_conv IN0): feat (the
// level's operator features, [N][kNetIn] BF16), u (FP32 scaled defect
// D^-1 d) and the np partial sums of ||u||^2 in part
struct LearnedInput {
  const uint16_t *feat;
  const float *u;
  const double *part;
  int np;
};
// the network's correction c on an M^3 lattice; nrm: ||u|| (read, or with
// `fused` written by the first layer); the input is net.act[0] unless fused
void net_forward_correction(LearnedNet &net, int M, double *nrm, double sqrt_count, float *c, hipStream_t st,
                            const LearnedInput *fused = nullptr);
// u = dinv d and the partials of ||u||^2 (learned_input_parts(n) of them)
void learned_input_sq(const float *dinv, const float *d, float *u, int64_t n, double *part, hipStream_t st);
int learned_input_parts(int64_t n);
// backward for the loss mean((out - t)^2) over interior nodes; accumulates into net.grads
// and returns (asynchronously) the loss into *loss_dev
void net_backward(LearnedNet &net, int M, const float *t, float *loss_dev, hipStream_t st);
// backward from a given FP32 gradient of the loss w.r.t. net.out (accumulates into net.grads)
void net_backward_from(LearnedNet &net, int M, const float *gy, hipStream_t st);
void net_zero_grads(LearnedNet &net, hipStream_t st);
// one Adam step with the accumulated gradient times gscale (1/B for the
// mean over a batch of B samples)
void net_adam(LearnedNet &net, float lr, hipStream_t st, float gscale = 1.0f);

struct DevLayout;
// dinv[p] = 1 / a_pp on interior nodes, 0 on Dirichlet nodes (FP32, lattice order)
void learned_diag_inv(const DevLayout &D, const float *val, float *dinv, hipStream_t st);
// u = dinv * d (lattice), the unscaled network input
void learned_input(const float *dinv, const float *d, float *u, int64_t n, hipStream_t st);
// x += alpha c with alpha = max(0, (d, ac) / (ac, ac)), in two launches
// (part: 2 kRedBlocks partials)
void learned_update(const float *d, const float *ac, const float *c, float *x, int64_t n, double *part,
                    hipStream_t st);
// learned_update with the defect d_in, then from the updated defect
// d = d_in - alpha ac (d_in may be d) the next step's u = dinv d (u may alias
// c) and the partials of ||u||^2 into pu (kRedBlocks doubles apart from
// part's 2 kRedBlocks), in two launches
void learned_update_carry(const float *d_in, float *d, const float *ac, const float *c, float *x, const float *dinv,
                          float *u, int64_t n, double *part, double *pu, hipStream_t st);
// net.act[0] = feat with channel 0 = bf16(u * sqrt_count / *nrm) (nrm: device
// FP64 2-norm of u; nrm == nullptr: no scaling)
void learned_build_input(LearnedNet &net, const float *u, const double *nrm, double sqrt_count, const uint16_t *feat,
                         int64_t n, hipStream_t st);
// operator features of a level: feat[p][1..6] = bf16(a(p, p +- e_k) / a(p,p)),
// feat[p][7..12] = bf16(a(p, p +- 2 e_k) / a(p,p)) on interior nodes (order x+,
// x-, y+, y-, z+, z-; 0 where the stencil has no such entry); channel 0, 13..15
// and boundary nodes 0
void learned_features(const DevLayout &D, const float *val, uint16_t *feat, hipStream_t st);
// out = in on interior nodes of an M^3 lattice, 0 on its boundary
void learned_mask(const float *in, int M, float *out, hipStream_t st);
// residual loss mean over interior nodes of (u - dinv * Ac)^2: its gradient
// w.r.t. Ac into g, partial sums into net.lpart (learned_loss_sum adds them up)
void learned_res_loss(LearnedNet &net, int M, const float *u, const float *dinv, const float *ac, float *g,
                      hipStream_t st);
void learned_loss_sum(LearnedNet &net, float *loss_dev, hipStream_t st);
// valT = the interior stencil operator's transpose in the CSS layout of val
// (slot d of row j: a_{j+d, j}; D.nvals entries)
void css_transpose_values(const DevLayout &D, const float *val, float *valT, hipStream_t st);
// y = transpose of the interior stencil operator applied to g (0 on the
// boundary), valT from css_transpose_values: a stencil product
void css_transpose(const DevLayout &D, const float *valT, const float *g, float *y, hipStream_t st);
// uniform(-1, 1) values on interior nodes of an M^3 lattice, 0 on the boundary
void learned_random(float *x, int M, uint32_t seed, hipStream_t st);
// t = sign * x * sqrt_count / *nrm on interior nodes, 0 on the boundary
void learned_target(const float *x, int M, const double *nrm, double sqrt_count, float sign, float *t,
                    hipStream_t st);

}  // namespace pmx
