// fdmamr_selector.hpp -- layout constants and launchers of the learned troubled-cell selector.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "fdmamr_device.hpp"

namespace fdm {

constexpr int MLP_IN = 80;  // 75 features (5x5 x {rho, p, |u|}) padded to a multiple of 16
constexpr int MLP_H = 64;

// FP32 master parameters: W1 [80][64] (rows 75..79 zero), b1, W2 [64][64], b2, W3, b3, w4, b4
constexpr int P_W1 = 0;
constexpr int P_B1 = P_W1 + MLP_IN * MLP_H;
constexpr int P_W2 = P_B1 + MLP_H;
constexpr int P_B2 = P_W2 + MLP_H * MLP_H;
constexpr int P_W3 = P_B2 + MLP_H;
constexpr int P_B3 = P_W3 + MLP_H * MLP_H;
constexpr int P_W4 = P_B3 + MLP_H;
constexpr int P_B4 = P_W4 + MLP_H;
constexpr int P_TOTAL = P_B4 + 1;            // 13569 (13249 live parameters)

// BF16 MFMA operand copies
constexpr int KW_W1T = 0;                     // [64][80]
constexpr int KW_W2T = KW_W1T + MLP_H * MLP_IN;
constexpr int KW_W3T = KW_W2T + MLP_H * MLP_H;
constexpr int KW_FWD = KW_W3T + MLP_H * MLP_H;  // forward needs the first three
constexpr int KW_W2N = KW_FWD;
constexpr int KW_W3N = KW_W2N + MLP_H * MLP_H;
constexpr int KW_PLAIN = KW_W3N + MLP_H * MLP_H;
// Inference operands in MFMA fragment order for v_mfma_f32_16x16x32_bf16: fragment f of a
// layer is 64 lanes x 8 BF16, lane-contiguous, so one 16-byte LDS read per lane fetches it.
// Layer 1: f = mt*3 + s (K = 96: the 75 features in the lane order the feature gather
// produces them); layers 2, 3: f = mt*2 + s.
constexpr int KF_FRAG = 64 * 8;
constexpr int KF_L1 = 0;
constexpr int KF_L2 = KF_L1 + 12 * KF_FRAG;
constexpr int KF_L3 = KF_L2 + 8 * KF_FRAG;
constexpr int KF_TOTAL = KF_L3 + 8 * KF_FRAG;   // 14336
constexpr int KW_FRAG = KW_PLAIN;
constexpr int KW_TOTAL = KW_FRAG + KF_TOTAL;

// FP32 vectors staged in LDS
constexpr int SB_B1 = 0, SB_B2 = 64, SB_B3 = 128, SB_W4 = 192, SB_B4 = 256, SB_TOTAL = 260;

// device counters: CNT_SLOTS copies of CNT_N counters (block b adds into slot b % CNT_SLOTS, so
// no address takes more than a few dozen atomics per step); readers sum the slots.
// CNT_TILES: 16-leaf tiles the inference ran the network on (the others were all near-uniform)
constexpr int CNT_TROUBLED = 0, CNT_JUDGED = 1, CNT_REJECTED = 2, CNT_TILES = 3, CNT_N = 4, CNT_SLOTS = 64;

struct SelectParams {
    float threshold;      // probability at or above which a leaf is troubled
    uint32_t judge_q16;   // judge sample fraction of cleared leaves, in 1/65536
    uint32_t step;        // step counter (varies the judge sample)
};

// inference: flags (and probabilities when prob != nullptr) of every interior leaf, in one
// persistent launch of two blocks per CU (n_cu: the device's compute units); nbr_beg, nbr, loc:
// the tile neighbourhoods of the tables (Tables::sel_nbr).
void launch_select(const DevTables& t, const float4* Qf, const int32_t* off16, const uint16_t* wpk,
                   const float* w32, const SelectParams& sp, uint8_t* flags, float* prob, int n_cu,
                   const int32_t* nbr_beg, const int32_t* nbr, const uint8_t* loc, int* counters,
                   hipStream_t st);
void launch_judge(int n, const float* label, float bound, uint8_t* flags, int* counters, hipStream_t st);
void launch_features(const DevTables& t, const float4* Qf, const int32_t* off16, uint16_t* X, int row0,
                     hipStream_t st);
void launch_labels_to_y(int n, const float* label, float bound, uint8_t* y, int row0, hipStream_t st);
void launch_mlp_grad(const uint16_t* wpk, const float* w32, const uint16_t* X, const uint8_t* y,
                     const int32_t* rows, int n, float pos_weight, float scale, float* gpart, float* lpart,
                     int nblocks, float* grad, float* loss, hipStream_t st);
void launch_adam(float* w, const float* grad, float* m, float* v, float lr, float b1, float b2, float eps,
                 float bc1, float bc2, hipStream_t st);
void launch_pack(const float* w, uint16_t* pk, hipStream_t st);

}  // namespace fdm
