/* Layouts of the neural fragment shader's network, shared by the host
 * library, the device kernels and shader modules.
 *
 * Network: 8 surface features -> Fourier encoding (32 frequencies per
 * feature, sine and cosine: 512 values) -> 64 -> ReLU -> 64 -> ReLU -> 3.
 * Feature d at frequency j (0..31) contributes, at encoding index
 * 64*d + 2*j + s, sin (s = 0) or cos (s = 1) of 2*pi*(j+1)*u_d, where
 * u_d = step_d * x_d and step_d (revolutions per unit of feature d) is a
 * property of the network, SRP_NN_FREQ_STEP unless set otherwise.
 *
 * Parameters (fp32 master copy, "canonical" layout, row = output):
 *   W1[64][512] b1[64] W2[64][64] b2[64] W3[3][64] b3[3]
 *
 * Matrix-core copies (bf16, written after every optimiser step).  The
 * layers are evaluated transposed (outputs x samples), so the result tile
 * of one layer is the right-hand operand of the next without leaving
 * registers; for that the K (input) order of the 64-wide layers is
 * permuted: K position p = 32*ks + 8*g + q of a 16x16x32 product step ks
 * holds hidden unit
 *     H(ks, g, q) = 16*(2*ks + q/4) + 4*g + q%4      (q = 0..7, g = 0..3).
 *   w1 [64][512]    W1 as is
 *   w2f[64][64]     w2f[m][p] = W2[m][H(p)]          (layer 2, forward)
 *   w3f[16][64]     w3f[c][p] = W3[c][H(p)], rows 3..15 zero
 *   w2b[64][64]     w2b[i][p] = W2[H(p)][i]          (layer 2 input gradient)
 *   w3b[64][32]     w3b[i][q] = W3[q][i] for q < 3, else zero
 */
#pragma once

#include <stdint.h>

#define SRP_NN_IN 8
#define SRP_NN_FREQ 32
#define SRP_NN_ENC 512
#define SRP_NN_HID 64
#define SRP_NN_OUT 3
#define SRP_NN_FREQ_STEP 0.5f

#define SRP_NN_OFF_W1 0
#define SRP_NN_OFF_B1 (SRP_NN_OFF_W1 + SRP_NN_HID * SRP_NN_ENC)
#define SRP_NN_OFF_W2 (SRP_NN_OFF_B1 + SRP_NN_HID)
#define SRP_NN_OFF_B2 (SRP_NN_OFF_W2 + SRP_NN_HID * SRP_NN_HID)
#define SRP_NN_OFF_W3 (SRP_NN_OFF_B2 + SRP_NN_HID)
#define SRP_NN_OFF_B3 (SRP_NN_OFF_W3 + SRP_NN_OUT * SRP_NN_HID)
#define SRP_NN_PARAMS (SRP_NN_OFF_B3 + SRP_NN_OUT)

/* Device view of a network (what a fragment shader's uniform points to). */
typedef struct SrpNnDevice
{
	const uint16_t* w1;    /* bf16 bits */
	const uint16_t* w2f;
	const uint16_t* w3f;
	const uint16_t* w2b;
	const uint16_t* w3b;
	const float* b1;
	const float* b2;
	const float* b3;
	float step[SRP_NN_IN]; /* frequency step of each feature's encoding */
} SrpNnDevice;
