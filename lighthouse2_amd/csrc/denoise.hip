/* denoise.hip - learned-denoiser kernels for gfx950 (see denoise.h).
 *
 * Convolutions are implicit GEMMs on the matrix cores,
 * v_mfma_f32_16x16x16_bf16 (bf16 operands, fp32 accumulation). Operand
 * layout of one 16x16x16 product in a wave of 64 lanes: A[m][k] with
 * m = lane % 16, k = 4 (lane / 16) + i; B[k][n] with n = lane % 16,
 * k = 4 (lane / 16) + i; D[m][n] with n = lane % 16, m = 4 (lane / 16) + i
 * (i = 0..3 the element of the lane's 4-vector).
 *
 * Forward / input gradient: M = 64 output pixels of one row per wave, N =
 * output channels, K = 9 taps x input channels; activations are NHWC, so the
 * 4 k-values a lane needs are 4 consecutive channels of one pixel (one 8-byte
 * load) and the weights [tap][n][k] give the same for B.
 * Weight gradient: M = output channels, N = input channels, K = pixels;
 * a 64-pixel tile of dy and the 3 x 66 pixel input halo are staged in LDS
 * channel-major so that both operands read consecutive pixels; each block
 * keeps partial sums for its tiles (the bias gradient from the same staged
 * rows) and a second kernel adds the partials in a fixed order
 * (deterministic).
 */
#include "denoise.h"

#include <mutex>

namespace lh {

typedef short s4v __attribute__((__vector_size__(4 * sizeof(short))));
typedef float f4v __attribute__((__vector_size__(4 * sizeof(float))));

__device__ __forceinline__ float bf2f(bf16_t b) { return __builtin_bit_cast(float, (uint32_t)b << 16); }
/* round to nearest even; NaN stays a quiet NaN */
__device__ __forceinline__ bf16_t f2bf(float f)
{
	const uint32_t u = __builtin_bit_cast(uint32_t, f);
	if ((u & 0x7fffffffu) > 0x7f800000u) return (bf16_t)((u >> 16) | 0x40u);
	return (bf16_t)((u + 0x7fffu + ((u >> 16) & 1u)) >> 16);
}

static const int DW_BLOCKS = 512;

/* Rows per block of the forward / input-gradient convolution's tiles on
 * large images (a wave per row, 64 pixels wide). */
static constexpr int TILE_ROWS = 8;

/* ---- 3x3 convolution, forward and input gradient ------------------------ */
template <int NB, bool RELU, bool F32OUT>
__global__ void __launch_bounds__(256) k_conv3x3(const bf16_t* __restrict__ x, int Cin, const bf16_t* __restrict__ w,
	const float* __restrict__ bias, void* __restrict__ y, int H, int W)
{
	const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
	const int row = blockIdx.y * 4 + wave;
	const int x0 = blockIdx.x * 64;
	if (row >= H) return;
	const int Cout = NB * 16;
	const int ml = lane & 15, kq = (lane >> 4) * 4;
	f4v acc[4][NB];
#pragma unroll
	for (int mb = 0; mb < 4; mb++)
#pragma unroll
		for (int nb = 0; nb < NB; nb++) acc[mb][nb] = (f4v){ 0.0f, 0.0f, 0.0f, 0.0f };
	for (int tap = 0; tap < 9; tap++)
	{
		const int yy = row + tap / 3 - 1, dx = tap % 3 - 1;
		const bool rowOk = yy >= 0 && yy < H;
		const bf16_t* xrow = x + (size_t)(rowOk ? yy : 0) * W * Cin;
		for (int c = 0; c < Cin; c += 16)
		{
			s4v b[NB];
#pragma unroll
			for (int nb = 0; nb < NB; nb++) b[nb] = *(const s4v*)(w + ((size_t)(tap * Cout + nb * 16 + ml) * Cin + c + kq));
#pragma unroll
			for (int mb = 0; mb < 4; mb++)
			{
				const int xx = x0 + mb * 16 + ml + dx;
				s4v a = (s4v){ 0, 0, 0, 0 };
				if (rowOk && xx >= 0 && xx < W) a = *(const s4v*)(xrow + (size_t)xx * Cin + c + kq);
#pragma unroll
				for (int nb = 0; nb < NB; nb++) acc[mb][nb] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b[nb], acc[mb][nb], 0, 0, 0);
			}
		}
	}
#pragma unroll
	for (int nb = 0; nb < NB; nb++)
	{
		const int ch = nb * 16 + ml;
		const float bv = bias ? bias[ch] : 0.0f;
#pragma unroll
		for (int mb = 0; mb < 4; mb++)
#pragma unroll
			for (int i = 0; i < 4; i++)
			{
				const int px = x0 + mb * 16 + kq + i;
				if (px >= W) continue;
				float v = acc[mb][nb][i] + bv;
				if (RELU) v = v > 0.0f ? v : 0.0f;
				const size_t o = ((size_t)row * W + px) * Cout + ch;
				if (F32OUT) ((float*)y)[o] = v;
				else ((bf16_t*)y)[o] = f2bf(v);
			}
	}
}

/* Tiled variant for the network's shapes: a block of ROWS waves computes
 * ROWS rows x 16 MBW pixels (one row per wave; 8 x 64 on large images, 4 x 64
 * with 64 input channels, 4 x 16 on small ones). The (ROWS + 2) x (16 MBW + 2)
 * pixel input halo is staged once
 * in LDS with 16-byte loads, all of a thread's loads in flight together, CP
 * channels per pixel (8 of padding against bank conflicts). Each wave then
 * runs the 9 taps as v_mfma_f32_16x16x32_bf16 products with
 * A = 16 output channels x 32 input channels of the [tap][Cout][Cin] weights
 * (read from memory one tap ahead of their use) and B = 32 input channels x
 * 16 pixels from LDS (lane: pixel lane % 16, channels 8 (lane / 16) + 0..7);
 * D then gives each lane 4 consecutive output channels of one pixel; bf16
 * outputs pair channel blocks across lanes for 16-byte writes (below).
 * CIN = 16 is padded to K = 32 with zero channels. CAT: the input is the
 * 64-channel concatenation [nearest 2x upsample of coarse (32 channels),
 * x (32 channels)]; the coarse pixels under the tile (and their halo) are staged
 * separately and upsampled by the fragment addressing. */
/* Decoders of the SVGF feature words (tools_shared.h:130-136,169-176):
 * albedo HDRtoRGB32 (10:11:11), normal PackNormal2, depth as float bits. */
__device__ __forceinline__ float3 feat_albedo(uint32_t c)
{
	return make_float3((float)(c >> 22) * (1.0f / 1023.0f), (float)((c >> 11) & 2047u) * (1.0f / 2047.0f), (float)(c & 2047u) * (1.0f / 2047.0f));
}
__device__ __forceinline__ float3 feat_normal(uint32_t p)
{
	return make_float3((float)((p >> 2) & 1023u) * (1.0f / 511.0f) - 1.0f, (float)((p >> 12) & 1023u) * (1.0f / 511.0f) - 1.0f,
		(float)(p >> 22) * (1.0f / 511.0f) - 1.0f);
}

static const float DN_ALBEDO_FLOOR = 0.02f;

/* The frame the output convolution's display variant (dn_conv3x3_display)
 * writes: features of the frame and its display image, w x h inside the
 * network's padded image. */
struct FrameIO
{
	const uint4* feat;
	int w, h;
	float4* image;
};

/* The display value of one pixel from the network's first 3 output channels,
 * as k_frame_output computes it. */
__device__ __forceinline__ float4 frame_display(float o0, float o1, float o2, uint32_t featX)
{
	const float3 a = feat_albedo(featX);
	const float o[3] = { o0, o1, o2 }, al[3] = { a.x, a.y, a.z };
	float v[3];
	for (int k = 0; k < 3; k++) v[k] = sqrtf(sqrtf(expm1f(fmaxf(o[k], 0.0f)) * fmaxf(al[k], DN_ALBEDO_FLOOR)));
	return make_float4(v[0], v[1], v[2], 0.0f);
}

typedef __bf16 bf8v __attribute__((ext_vector_type(8)));
enum { IO_NONE = 0, IO_DISPLAY = 1 };
template <int CIN, int NB, bool RELU, bool F32OUT, bool CAT, bool POOL = false, int MBW = 4, int ROWS = 4, int IO = IO_NONE>
__global__ void __launch_bounds__(64 * ROWS) k_conv3x3_t(const bf16_t* __restrict__ x, const bf16_t* __restrict__ coarse,
	const bf16_t* __restrict__ w, const float* __restrict__ bias, void* __restrict__ y, int H, int W, bf16_t* __restrict__ ypool,
	FrameIO io = FrameIO{})
{
	static_assert(IO != IO_DISPLAY || (NB == 1 && F32OUT && !RELU), "the display epilogue reads the first output channels");
	constexpr int CX = CAT ? 32 : CIN;                  /* channels of x */
	constexpr int CKX = CX < 32 ? 32 : CX;              /* channels per pixel of the x tile (K, padded) */
	constexpr int CP = CKX + 8, TW = 16 * MBW + 2, TR = ROWS + 2, CH8 = CX / 8, CHP = CKX / 8;
	constexpr int KS = CAT ? 2 : CKX / 32;              /* 32-channel steps of K per tap */
	constexpr int WCIN = CAT ? 64 : CIN;                /* weight row length */
	constexpr int CPC = 40, TWC = 8 * MBW + 2, TRC = ROWS / 2 + 2; /* coarse tile */
	constexpr int NT = 64 * ROWS;                       /* threads per block: a wave per output row */
	constexpr int Cout = NB * 16;
	__shared__ __attribute__((aligned(16))) bf16_t xs[TR * TW * CP];
	__shared__ __attribute__((aligned(16))) bf16_t cs[CAT ? TRC * TWC * CPC : 8];
	const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6, ml = lane & 15, kg = lane >> 4;
	const int tx = blockIdx.x, ty = blockIdx.y;
	const int x0 = tx * (16 * MBW), r0 = ty * ROWS;
	{
		constexpr int N = TR * TW * CHP, IT = (N + NT - 1) / NT, BATCH = 8;
#pragma unroll
		for (int it0 = 0; it0 < IT; it0 += BATCH)
		{
			uint4 v[BATCH];
#pragma unroll
			for (int j = 0; j < BATCH; j++)
			{
				const int i = tid + (it0 + j) * NT;
				const int c8 = i % CHP, p = (i / CHP) % TW, gy = r0 - 1 + i / (CHP * TW), gx = x0 - 1 + p;
				v[j] = it0 + j < IT && i < N && c8 < CH8 && gy >= 0 && gy < H && gx >= 0 && gx < W ?
					*(const uint4*)(x + ((size_t)gy * W + gx) * CX + c8 * 8) : make_uint4(0, 0, 0, 0);
			}
#pragma unroll
			for (int j = 0; j < BATCH; j++)
			{
				const int i = tid + (it0 + j) * NT;
				if (it0 + j < IT && i < N) *(uint4*)(xs + (i / CHP) * CP + (i % CHP) * 8) = v[j];
			}
		}
		if (CAT)
		{
			constexpr int NC = TRC * TWC * 4, ITC = (NC + NT - 1) / NT;
			const int Hc = H >> 1, Wc = W >> 1;
			uint4 v[ITC];
#pragma unroll
			for (int j = 0; j < ITC; j++)
			{
				const int i = tid + j * NT;
				const int c8 = i % 4, p = (i / 4) % TWC, cy = (r0 >> 1) - 1 + i / (4 * TWC), cx = (x0 >> 1) - 1 + p;
				v[j] = i < NC && cy >= 0 && cy < Hc && cx >= 0 && cx < Wc ? *(const uint4*)(coarse + ((size_t)cy * Wc + cx) * 32 + c8 * 8) : make_uint4(0, 0, 0, 0);
			}
#pragma unroll
			for (int j = 0; j < ITC; j++)
			{
				const int i = tid + j * NT;
				if (i < NC) *(uint4*)(cs + (i / 4) * CPC + (i % 4) * 8) = v[j];
			}
		}
	}
	auto wfrag = [&](int tap, int ks, int nb) -> bf8v {
		const int k = ks * 32 + kg * 8;
		return k < WCIN ? *(const bf8v*)(w + ((size_t)(tap * Cout + nb * 16 + ml) * WCIN + k)) : (bf8v){ 0, 0, 0, 0, 0, 0, 0, 0 };
	};
	/* bias now, beside the tile's loads, so that the epilogue waits on nothing;
	 * weight fragments two taps ahead of their products */
	float bv[NB][4];
#pragma unroll
	for (int nb = 0; nb < NB; nb++)
#pragma unroll
		for (int i = 0; i < 4; i++) bv[nb][i] = bias ? bias[nb * 16 + kg * 4 + i] : 0.0f;
	/* weight fragments PD taps ahead of their products: two where the third
	 * set of registers costs no occupancy (64-pixel tiles, 32 input channels,
	 * up to 32 outputs), else one with the tap loop kept rolled */
	constexpr int PD = MBW > 1 && (CIN == 32 || CAT) && NB <= 2 ? 2 : 1;
	bf8v wc[KS][NB], w1[KS][NB], w2[KS][NB];
#pragma unroll
	for (int ks = 0; ks < KS; ks++)
#pragma unroll
		for (int nb = 0; nb < NB; nb++)
		{
			wc[ks][nb] = wfrag(0, ks, nb);
			if (PD == 2) w1[ks][nb] = wfrag(1, ks, nb);
		}
	f4v acc[MBW][NB];
#pragma unroll
	for (int mb = 0; mb < MBW; mb++)
#pragma unroll
		for (int nb = 0; nb < NB; nb++) acc[mb][nb] = (f4v){ 0.0f, 0.0f, 0.0f, 0.0f };
	__syncthreads();
	const auto products = [&](int tap) {
		const int dy = tap / 3, dx = tap % 3;
#pragma unroll
		for (int ks = 0; ks < KS; ks++)
#pragma unroll
			for (int mb = 0; mb < MBW; mb++)
			{
				const int m = mb * 16 + ml;
				bf8v a;
				if (CAT && ks == 0) a = *(const bf8v*)(cs + ((((wave + dy - 1) >> 1) + 1) * TWC + ((m + dx - 1) >> 1) + 1) * CPC + kg * 8);
				else a = *(const bf8v*)(xs + ((wave + dy) * TW + m + dx) * CP + (CAT ? 0 : ks * 32) + kg * 8);
#pragma unroll
				for (int nb = 0; nb < NB; nb++) acc[mb][nb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(wc[ks][nb], a, acc[mb][nb], 0, 0, 0);
			}
	};
	if constexpr (PD == 2)
	{
#pragma unroll
		for (int tap = 0; tap < 9; tap++)
		{
			if (tap < 7)
#pragma unroll
				for (int ks = 0; ks < KS; ks++)
#pragma unroll
					for (int nb = 0; nb < NB; nb++) w2[ks][nb] = wfrag(tap + 2, ks, nb);
			products(tap);
#pragma unroll
			for (int ks = 0; ks < KS; ks++)
#pragma unroll
				for (int nb = 0; nb < NB; nb++) wc[ks][nb] = w1[ks][nb], w1[ks][nb] = w2[ks][nb];
		}
	}
	else
	{
#pragma unroll 1
		for (int tap = 0; tap < 9; tap++)
		{
			if (tap < 8)
#pragma unroll
				for (int ks = 0; ks < KS; ks++)
#pragma unroll
					for (int nb = 0; nb < NB; nb++) w1[ks][nb] = wfrag(tap + 1, ks, nb);
			products(tap);
#pragma unroll
			for (int ks = 0; ks < KS; ks++)
#pragma unroll
				for (int nb = 0; nb < NB; nb++) wc[ks][nb] = w1[ks][nb];
		}
	}
	const int ry = r0 + wave;
	float hs[POOL ? MBW : 1][NB][4];                     /* POOL: sums of horizontal pixel pairs */
	/* bf16 output with channel blocks in pairs: lanes kg and kg ^ 1 swap half
	 * their fragments so that each lane writes 8 consecutive channels (16 B)
	 * and one store instruction covers whole 64 B pixel rows; kg even keeps
	 * block nb (channels 4 kg .. 4 kg + 7), kg odd block nb + 1 (channels
	 * 16 + 4 (kg - 1) .. + 7). */
	constexpr bool PAIRED = !F32OUT && (NB % 2) == 0;
	uint32_t pk[MBW][NB][2];
#pragma unroll
	for (int nb = 0; nb < NB; nb++)
	{
		const int ch = nb * 16 + kg * 4;
#pragma unroll
		for (int mb = 0; mb < MBW; mb++)
		{
			const int px = x0 + mb * 16 + ml;
			float v[4];
#pragma unroll
			for (int i = 0; i < 4; i++)
			{
				v[i] = acc[mb][nb][i] + bv[nb][i];
				if (RELU) v[i] = v[i] > 0.0f ? v[i] : 0.0f;
			}
			const size_t o = ((size_t)ry * W + px) * Cout + ch;
			if (PAIRED)
				pk[mb][nb][0] = (uint32_t)f2bf(v[0]) | (uint32_t)f2bf(v[1]) << 16, pk[mb][nb][1] = (uint32_t)f2bf(v[2]) | (uint32_t)f2bf(v[3]) << 16;
			else if (IO == IO_DISPLAY)
			{	/* lanes holding channels 0..3 write the display pixel of the frame */
				if (kg == 0 && ry < io.h && px < io.w)
				{
					const size_t q = (size_t)px + (size_t)ry * io.w;
					io.image[q] = frame_display(v[0], v[1], v[2], io.feat[q].x);
				}
			}
			else if (ry < H && px < W)
			{
				if (F32OUT) *(float4*)((float*)y + o) = make_float4(v[0], v[1], v[2], v[3]);
				else *(uint2*)((bf16_t*)y + o) = make_uint2((uint32_t)f2bf(v[0]) | (uint32_t)f2bf(v[1]) << 16,
					(uint32_t)f2bf(v[2]) | (uint32_t)f2bf(v[3]) << 16);
			}
			if (POOL)                                     /* the stored (bf16) values, left + right */
#pragma unroll
				for (int i = 0; i < 4; i++)
				{
					const float r = bf2f(f2bf(v[i]));
					hs[POOL ? mb : 0][nb][i] = r + __shfl_xor(r, 1);
				}
		}
	}
	if (PAIRED)
	{
		const bool even = !(kg & 1);
#pragma unroll
		for (int mb = 0; mb < MBW; mb++)
		{
			const int px = x0 + mb * 16 + ml;
#pragma unroll
			for (int nb = 0; nb < NB; nb += 2)
			{
				const uint32_t s0 = even ? pk[mb][nb + 1][0] : pk[mb][nb][0], s1 = even ? pk[mb][nb + 1][1] : pk[mb][nb][1];
				const uint32_t r0 = (uint32_t)__shfl_xor((int)s0, 16), r1 = (uint32_t)__shfl_xor((int)s1, 16);
				const uint4 q = even ? make_uint4(pk[mb][nb][0], pk[mb][nb][1], r0, r1) : make_uint4(r0, r1, pk[mb][nb + 1][0], pk[mb][nb + 1][1]);
				const int ch = even ? nb * 16 + kg * 4 : (nb + 1) * 16 + (kg - 1) * 4;
				if (ry < H && px < W) *(uint4*)((bf16_t*)y + ((size_t)ry * W + px) * Cout + ch) = q;
			}
		}
	}
	if (POOL)
	{
		/* 2x2 average: (top pair) + (bottom pair), as dn_pool2 on the stored tensor.
		 * Odd rows pass their pair sums to the even rows through LDS. */
		float* ex = (float*)xs;
		__syncthreads();                                  /* every wave is done with the staged tile */
		if (wave & 1)
#pragma unroll
			for (int mb = 0; mb < MBW; mb++)
#pragma unroll
				for (int nb = 0; nb < NB; nb++)
#pragma unroll
					for (int i = 0; i < 4; i++) ex[(((wave >> 1) * MBW + mb) * NB * 4 + nb * 4 + i) * 64 + lane] = hs[POOL ? mb : 0][nb][i];
		__syncthreads();
		const int Wp = W >> 1, pr = (r0 >> 1) + (wave >> 1);
		if (!(wave & 1) && ry + 1 < H && !(ml & 1))
#pragma unroll
			for (int mb = 0; mb < MBW; mb++)
			{
				const int px = x0 + mb * 16 + ml;
				if (px >= W) continue;
#pragma unroll
				for (int nb = 0; nb < NB; nb++)
				{
					bf16_t q[4];
#pragma unroll
					for (int i = 0; i < 4; i++)
						q[i] = f2bf((hs[POOL ? mb : 0][nb][i] + ex[(((wave >> 1) * MBW + mb) * NB * 4 + nb * 4 + i) * 64 + lane]) * 0.25f);
					*(uint2*)(ypool + ((size_t)pr * Wp + (px >> 1)) * Cout + nb * 16 + kg * 4) =
						make_uint2((uint32_t)q[0] | (uint32_t)q[1] << 16, (uint32_t)q[2] | (uint32_t)q[3] << 16);
				}
			}
	}
}

/* Images with fewer than SMALL_TILES 64-pixel-wide tiles of 4 rows (levels
 * 3 and 4 of the network at 1080p) run 4 x 16 tiles: four times the blocks,
 * so that a small layer spreads over more compute units instead of running
 * at the latency of a few blocks. (At level 2, 4 x 16 tiles were slower: a
 * block's nine dependent weight loads then outweigh its products.) */
static const int SMALL_TILES = 256;
static bool small_image(int H, int W) { return ((W + 63) / 64) * ((H + 3) / 4) < SMALL_TILES; }


template <int CIN, int NB, bool CAT, int MBW, int ROWS>
static void conv_t_w(const bf16_t* x, const bf16_t* coarse, const bf16_t* w, const float* bias, void* y, int H, int W, int flags, hipStream_t s)
{
	const dim3 grid((W + 16 * MBW - 1) / (16 * MBW), (H + ROWS - 1) / ROWS), block(64 * ROWS);
	const bool relu = flags & DN_RELU, f32 = flags & DN_OUT_F32;
	if (relu && f32) k_conv3x3_t<CIN, NB, true, true, CAT, false, MBW, ROWS><<<grid, block, 0, s>>>(x, coarse, w, bias, y, H, W, nullptr);
	else if (relu) k_conv3x3_t<CIN, NB, true, false, CAT, false, MBW, ROWS><<<grid, block, 0, s>>>(x, coarse, w, bias, y, H, W, nullptr);
	else if (f32) k_conv3x3_t<CIN, NB, false, true, CAT, false, MBW, ROWS><<<grid, block, 0, s>>>(x, coarse, w, bias, y, H, W, nullptr);
	else k_conv3x3_t<CIN, NB, false, false, CAT, false, MBW, ROWS><<<grid, block, 0, s>>>(x, coarse, w, bias, y, H, W, nullptr);
}

template <int CIN, int NB, bool CAT>
static hipError_t conv_t(const bf16_t* x, const bf16_t* coarse, const bf16_t* w, const float* bias, void* y, int H, int W, int flags, hipStream_t s)
{
	/* 64 input channels without the concatenation keep 4-row tiles (an 8-row tile would not leave room for two blocks per CU) */
	constexpr int ROWS = (CIN == 64 && !CAT) ? 4 : TILE_ROWS;
	if (small_image(H, W)) conv_t_w<CIN, NB, CAT, 1, 4>(x, coarse, w, bias, y, H, W, flags, s);
	else conv_t_w<CIN, NB, CAT, 4, ROWS>(x, coarse, w, bias, y, H, W, flags, s);
	return hipGetLastError();
}

hipError_t dn_conv3x3_pool(const bf16_t* x, int Cin, const bf16_t* w, const float* bias, bf16_t* y, int Cout, int H, int W,
	bf16_t* ypool, hipStream_t s)
{
	if (Cout != 32 || (Cin != 16 && Cin != 32) || H <= 0 || W <= 0 || H % 2 || W % 2) return hipErrorInvalidValue;
	const dim3 block(256);
	if (small_image(H, W))
	{
		const dim3 grid((W + 15) / 16, (H + 3) / 4);
		if (Cin == 16) k_conv3x3_t<16, 2, true, false, false, true, 1><<<grid, block, 0, s>>>(x, nullptr, w, bias, y, H, W, ypool);
		else k_conv3x3_t<32, 2, true, false, false, true, 1><<<grid, block, 0, s>>>(x, nullptr, w, bias, y, H, W, ypool);
	}
	else
	{
		const dim3 grid((W + 63) / 64, (H + TILE_ROWS - 1) / TILE_ROWS), blockL(64 * TILE_ROWS);
		if (Cin == 16) k_conv3x3_t<16, 2, true, false, false, true, 4, TILE_ROWS><<<grid, blockL, 0, s>>>(x, nullptr, w, bias, y, H, W, ypool);
		else k_conv3x3_t<32, 2, true, false, false, true, 4, TILE_ROWS><<<grid, blockL, 0, s>>>(x, nullptr, w, bias, y, H, W, ypool);
	}
	return hipGetLastError();
}

hipError_t dn_conv3x3_cat(const bf16_t* coarse, const bf16_t* skip, const bf16_t* w, const float* bias, void* y, int Cout, int H, int W,
	int flags, hipStream_t s)
{
	if (Cout != 32 || H <= 0 || W <= 0 || H % 2 || W % 2) return hipErrorInvalidValue;
	return conv_t<64, 2, true>(skip, coarse, w, bias, y, H, W, flags, s);
}

template <int NB>
static hipError_t conv_nb(const bf16_t* x, int Cin, const bf16_t* w, const float* bias, void* y, int H, int W, int flags, hipStream_t s)
{
	const dim3 grid((W + 63) / 64, (H + 3) / 4), block(256);
	const bool relu = flags & DN_RELU, f32 = flags & DN_OUT_F32;
	if (relu && f32) k_conv3x3<NB, true, true><<<grid, block, 0, s>>>(x, Cin, w, bias, y, H, W);
	else if (relu) k_conv3x3<NB, true, false><<<grid, block, 0, s>>>(x, Cin, w, bias, y, H, W);
	else if (f32) k_conv3x3<NB, false, true><<<grid, block, 0, s>>>(x, Cin, w, bias, y, H, W);
	else k_conv3x3<NB, false, false><<<grid, block, 0, s>>>(x, Cin, w, bias, y, H, W);
	return hipGetLastError();
}

hipError_t dn_conv3x3(const bf16_t* x, int Cin, const bf16_t* w, const float* bias, void* y, int Cout, int H, int W, int flags, hipStream_t s)
{
	if (Cin <= 0 || Cin % 16 || Cout <= 0 || Cout % 16 || Cout > 64 || H <= 0 || W <= 0) return hipErrorInvalidValue;
	if (Cin == 16 && Cout == 32) return conv_t<16, 2, false>(x, nullptr, w, bias, y, H, W, flags, s);
	if (Cin == 16 && Cout == 16) return conv_t<16, 1, false>(x, nullptr, w, bias, y, H, W, flags, s);
	if (Cin == 32)
	{
		if (Cout == 16) return conv_t<32, 1, false>(x, nullptr, w, bias, y, H, W, flags, s);
		if (Cout == 32) return conv_t<32, 2, false>(x, nullptr, w, bias, y, H, W, flags, s);
		if (Cout == 64) return conv_t<32, 4, false>(x, nullptr, w, bias, y, H, W, flags, s);
	}
	if (Cin == 64 && Cout == 32) return conv_t<64, 2, false>(x, nullptr, w, bias, y, H, W, flags, s);
	switch (Cout / 16)
	{
	case 1: return conv_nb<1>(x, Cin, w, bias, y, H, W, flags, s);
	case 2: return conv_nb<2>(x, Cin, w, bias, y, H, W, flags, s);
	case 3: return conv_nb<3>(x, Cin, w, bias, y, H, W, flags, s);
	default: return conv_nb<4>(x, Cin, w, bias, y, H, W, flags, s);
	}
}

/* The output convolution (32 -> 16) writing the display image of a
 * filter-core frame from its epilogue: equal to dn_conv3x3 (fp32 out) +
 * dn_frame_output bit for bit, without the fp32 output tensor. (Reading the
 * frame into the first convolution's tile the same way was measured slower
 * than dn_frame_input + dn_conv3x3: 181 against 99 µs at 1080p.) */
template <int CIN, int NB, bool RELU, bool F32OUT, int IO>
static void conv_io(const bf16_t* x, const bf16_t* w, const float* bias, void* y, int H, int W, const FrameIO& io, hipStream_t s)
{
	if (small_image(H, W))
		k_conv3x3_t<CIN, NB, RELU, F32OUT, false, false, 1, 4, IO><<<dim3((W + 15) / 16, (H + 3) / 4), 256, 0, s>>>(x, nullptr, w, bias, y, H, W,
			nullptr, io);
	else
		k_conv3x3_t<CIN, NB, RELU, F32OUT, false, false, 4, TILE_ROWS, IO><<<dim3((W + 63) / 64, (H + TILE_ROWS - 1) / TILE_ROWS), 64 * TILE_ROWS, 0,
			s>>>(x, nullptr, w, bias, y, H, W, nullptr, io);
}

hipError_t dn_conv3x3_display(const bf16_t* x, const bf16_t* wts, const float* bias, const uint4* features, int w, int h, float4* image,
	int Hp, int Wp, hipStream_t s)
{
	if (w <= 0 || h <= 0 || Hp < h || Wp < w) return hipErrorInvalidValue;
	const FrameIO io{ features, w, h, image };
	conv_io<32, 1, false, true, IO_DISPLAY>(x, wts, bias, nullptr, Hp, Wp, io, s);
	return hipGetLastError();
}

/* ---- weight gradient ---------------------------------------------------
 * dW[tap][co][ci] = sum_p dy[p][co] x[p + tap - (1,1)][ci] as a GEMM over
 * pixels on v_mfma_f32_16x16x32_bf16 (M = output channels, N = input
 * channels, K = 32 pixels). Writing q = p + (dx - 1) for the input pixel, a
 * tap pairs input row ty + r - 1 at pixels q with the output-gradient row ty
 * at pixels q + 1 - dx, so a block keeps, per 64-pixel segment of a row:
 * the three input rows (a ring: one new row per step) and the gradient row
 * at its three horizontal shifts, all channel-major ([channel][pixel]) so
 * that every MFMA fragment (8 consecutive pixels of one channel) is one
 * aligned 16-byte LDS read. Staging lanes load 4 pixels x 8 channels with
 * 16-byte global loads and write them transposed as 8 rows of 4 pixels; the
 * next row's loads are in flight while the current row's products run. Nine waves, one per tap, accumulate all Cout x Cin tiles of
 * their tap in registers over a column segment of rows; per-block partial
 * sums are reduced in block order (deterministic). */
typedef __bf16 bf8v_dw __attribute__((ext_vector_type(8)));

/* Channel c of four pixels: v[p] holds channels 0..7 of pixel p (bf16
 * pairs per dword); the result holds pixels 0..3 of channel c. */
__device__ __forceinline__ uint2 column4(const uint4 (&v)[4], int c)
{
	uint32_t d[4];
#pragma unroll
	for (int p = 0; p < 4; p++)
		d[p] = (c >> 1) == 0 ? v[p].x : (c >> 1) == 1 ? v[p].y : (c >> 1) == 2 ? v[p].z : v[p].w;
	if (c & 1) return make_uint2((d[0] >> 16) | (d[1] & 0xffff0000u), (d[2] >> 16) | (d[3] & 0xffff0000u));
	return make_uint2((d[0] & 0xffffu) | (d[1] << 16), (d[2] & 0xffffu) | (d[3] << 16));
}

template <int COUT>
__device__ __forceinline__ bool tid_in_bias(int tid) { return tid < 16 * COUT; }

template <int CIN, int COUT>
__global__ void __launch_bounds__(576) k_conv3x3_dw(const bf16_t* __restrict__ x, const bf16_t* __restrict__ dy, int H, int W,
	int rowsPerChunk, float* __restrict__ partial, float* __restrict__ partialB)
{
	constexpr int MB = COUT / 16, NB = CIN / 16, KP = 64 + 8;
	constexpr int XL = 2 * CIN, DL = 6 * COUT;   /* staging lanes: 16 groups of 4 pixels x CIN / 8 channel groups; 3 shifts x 16 x COUT / 8 */
	__shared__ __attribute__((aligned(16))) bf16_t xs[3][CIN][KP];
	__shared__ __attribute__((aligned(16))) bf16_t ds[3][COUT][KP];
	__shared__ float bred[16][COUT];
	/* bias gradient: lane (channel tid % COUT, pixel group tid / COUT) adds 4 pixels of each row */
	const bool isB = partialB != nullptr && tid_in_bias<COUT>(threadIdx.x);
	float bsum = 0.0f;
	const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
	const int r = wave / 3, dxo = wave % 3, sh = 2 - dxo;
	const int tilesX = (W + 63) / 64;
	const int x0 = (int)(blockIdx.x % tilesX) * 64, y0 = (int)(blockIdx.x / tilesX) * rowsPerChunk;
	const int y1 = y0 + rowsPerChunk < H ? y0 + rowsPerChunk : H;
	/* staging role: x (tid < XL) or dy (XL <= tid < XL + DL) */
	const bool isX = tid < XL, isD = !isX && tid < XL + DL;
	const int t = isX ? tid : tid - XL;
	const int cg = isX ? t % (CIN / 8) : t % (COUT / 8);
	const int q = isX ? t / (CIN / 8) : (t / (COUT / 8)) % 16;
	const int shift = isX ? 0 : t / (2 * COUT);    /* dy copy s holds pixels x0 + j + s - 1 */
	uint4 v[4];
	const auto load = [&](int row) {
		const bf16_t* base = isX ? x : dy;
		const int C = isX ? CIN : COUT;
		const bool rowOk = row >= 0 && row < H;
#pragma unroll
		for (int p = 0; p < 4; p++)
		{
			const int px = x0 + 4 * q + p + (isX ? 0 : shift - 1);
			v[p] = (rowOk && px >= 0 && px < W) ? *(const uint4*)(base + ((size_t)row * W + px) * C + cg * 8) : make_uint4(0, 0, 0, 0);
		}
	};
	const auto store = [&](int slot) {
#pragma unroll
		for (int c = 0; c < 8; c++)
		{
			bf16_t* dst = isX ? &xs[slot][cg * 8 + c][4 * q] : &ds[shift][cg * 8 + c][4 * q];
			*(uint2*)dst = column4(v, c);
		}
	};
	const auto slotOf = [](int row) { return (row % 3 + 3) % 3; };
	f4v acc[MB][NB];
#pragma unroll
	for (int mb = 0; mb < MB; mb++)
#pragma unroll
		for (int nb = 0; nb < NB; nb++) acc[mb][nb] = (f4v){ 0.0f, 0.0f, 0.0f, 0.0f };
	if (y0 < y1)
	{
		if (isX)
		{	/* the two rows above the first step's new row */
			load(y0 - 1), store(slotOf(y0 - 1));
			load(y0), store(slotOf(y0));
		}
		if (isX) load(y0 + 1);
		else if (isD) load(y0);
		for (int ty = y0; ty < y1; ty++)
		{
			if (isX) store(slotOf(ty + 1));
			else if (isD) store(0);
			__syncthreads();
			if (ty + 1 < y1)
			{	/* the next step's rows, in flight during this step's products */
				if (isX) load(ty + 2);
				else if (isD) load(ty + 1);
			}
			if (isB)
			{	/* the unshifted copy holds row ty at pixels x0 .. x0 + 63 (zero beyond W) */
				const uint2 v = *(const uint2*)&ds[1][tid % COUT][(tid / COUT) * 4];
				bsum += bf2f((bf16_t)(v.x & 0xffffu)), bsum += bf2f((bf16_t)(v.x >> 16));
				bsum += bf2f((bf16_t)(v.y & 0xffffu)), bsum += bf2f((bf16_t)(v.y >> 16));
			}
			const int xslot = slotOf(ty - 1 + r);
#pragma unroll
			for (int kc = 0; kc < 2; kc++)
			{
				const int koff = kc * 32 + (lane >> 4) * 8;
				bf8v_dw A[MB], B[NB];
#pragma unroll
				for (int mb = 0; mb < MB; mb++) A[mb] = *(const bf8v_dw*)&ds[sh][mb * 16 + (lane & 15)][koff];
#pragma unroll
				for (int nb = 0; nb < NB; nb++) B[nb] = *(const bf8v_dw*)&xs[xslot][nb * 16 + (lane & 15)][koff];
#pragma unroll
				for (int mb = 0; mb < MB; mb++)
#pragma unroll
					for (int nb = 0; nb < NB; nb++) acc[mb][nb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A[mb], B[nb], acc[mb][nb], 0, 0, 0);
			}
			__syncthreads();
		}
	}
	if (partialB)
	{	/* this block's bias partial: the 16 pixel groups of each channel added in group order */
		if (isB) bred[tid / COUT][tid % COUT] = bsum;
		__syncthreads();
		if (tid < COUT)
		{
			float t = 0.0f;
			for (int g = 0; g < 16; g++) t += bred[g][tid];
			partialB[(size_t)blockIdx.x * COUT + tid] = t;
		}
	}
	/* D layout: lane holds rows 4 (lane / 16) + i, column lane % 16 */
	float* out = partial + (size_t)blockIdx.x * 9 * COUT * CIN;
	const int tap = wave;
#pragma unroll
	for (int mb = 0; mb < MB; mb++)
#pragma unroll
		for (int nb = 0; nb < NB; nb++)
#pragma unroll
			for (int i = 0; i < 4; i++)
				out[((size_t)tap * COUT + mb * 16 + (lane >> 4) * 4 + i) * CIN + nb * 16 + (lane & 15)] = acc[mb][nb][i];
}

/* out[i] = sum over b < blocks of partial[b][i] in a fixed order: 8 lanes
 * per output each add every 8th block partial (32 consecutive outputs per
 * block row, coalesced), then the 8 sums are added in lane order. */
__global__ void __launch_bounds__(256) k_reduce_partials(const float* __restrict__ partial, int blocks, int n, float* __restrict__ out)
{
	__shared__ float red[8][33];
	const int lane = threadIdx.x & 31, part = threadIdx.x >> 5;
	const int i = blockIdx.x * 32 + lane;
	float s = 0.0f;
	if (i < n)
		for (int b = part; b < blocks; b += 8) s += partial[(size_t)b * n + i];
	red[part][lane] = s;
	__syncthreads();
	if (part == 0 && i < n)
	{
		float t = 0.0f;
#pragma unroll
		for (int k = 0; k < 8; k++) t += red[k][lane];
		out[i] = t;
	}
}

size_t dn_dw_workspace_bytes(int Cin, int Cout) { return (size_t)DW_BLOCKS * (9 * Cout * Cin + Cout) * sizeof(float); }

template <int CIN, int COUT>
static void launch_dw(const bf16_t* x, const bf16_t* dy, int H, int W, int rowsPerChunk, float* partial, float* partialB, int blocks, hipStream_t s)
{
	k_conv3x3_dw<CIN, COUT><<<blocks, 576, 0, s>>>(x, dy, H, W, rowsPerChunk, partial, partialB);
}

hipError_t dn_conv3x3_dw(const bf16_t* x, int Cin, const bf16_t* dy, int Cout, int H, int W, float* dw, float* db, void* ws, hipStream_t s)
{
	if (H <= 0 || W <= 0 || !ws) return hipErrorInvalidValue;
	/* blocks = 64-pixel column segments x row chunks (at least 8 rows each), at most DW_BLOCKS */
	const int tilesX = (W + 63) / 64;
	int chunks = DW_BLOCKS / tilesX;
	if (chunks < 1) chunks = 1;
	if (chunks > (H + 7) / 8) chunks = (H + 7) / 8;
	const int rowsPerChunk = (H + chunks - 1) / chunks;
	const int blocks = tilesX * ((H + rowsPerChunk - 1) / rowsPerChunk);
	if (blocks > DW_BLOCKS) return hipErrorInvalidValue;   /* W beyond DW_BLOCKS * 64 pixels */
	float* partial = (float*)ws;
	const int n = 9 * Cout * Cin;
	float* bp = db ? partial + (size_t)DW_BLOCKS * n : nullptr;   /* bias partials of the same blocks */
	if (Cin == 16 && Cout == 32) launch_dw<16, 32>(x, dy, H, W, rowsPerChunk, partial, bp, blocks, s);
	else if (Cin == 32 && Cout == 32) launch_dw<32, 32>(x, dy, H, W, rowsPerChunk, partial, bp, blocks, s);
	else if (Cin == 64 && Cout == 32) launch_dw<64, 32>(x, dy, H, W, rowsPerChunk, partial, bp, blocks, s);
	else if (Cin == 32 && Cout == 16) launch_dw<32, 16>(x, dy, H, W, rowsPerChunk, partial, bp, blocks, s);
	else if (Cin == 16 && Cout == 16) launch_dw<16, 16>(x, dy, H, W, rowsPerChunk, partial, bp, blocks, s);
	else return hipErrorInvalidValue;
	k_reduce_partials<<<(n + 31) / 32, 256, 0, s>>>(partial, blocks, n, dw);
	if (db) k_reduce_partials<<<(Cout + 31) / 32, 256, 0, s>>>(bp, blocks, Cout, db);
	return hipGetLastError();
}

/* ---- weight gradient without allocation --------------------------------
 * The workspace of dn_conv3x3_dw_noalloc is a file-scope device array sized
 * for the largest supported shape (Cin 64, Cout 32: DW_BLOCKS x (9 x 32 x 64
 * + 32) floats, 37.8 MB); each device holds its own copy until the process
 * ends. Every block partial that the reduction reads is written by the same
 * call first, so nothing needs clearing between calls; calls that overlap
 * on one device share the array and race. */
static constexpr size_t DW_SCRATCH_FLOATS = (size_t)DW_BLOCKS * (9 * 32 * 64 + 32);
static __device__ __align__(256) float scratch[DW_SCRATCH_FLOATS];

static constexpr int MAX_DEVICES = 64;
static std::mutex g_scratchLock;
static void* g_scratchPtr[MAX_DEVICES];   /* device address of scratch per device, looked up once */

size_t dn_dw_scratch_bytes() { return DW_SCRATCH_FLOATS * sizeof(float); }

hipError_t dn_conv3x3_dw_noalloc(const bf16_t* x, int Cin, const bf16_t* dy, int Cout, int H, int W, float* dw, float* db, hipStream_t s)
{
	if (dn_dw_workspace_bytes(Cin, Cout) > dn_dw_scratch_bytes()) return hipErrorInvalidValue;
	int dev = 0;
	hipError_t e = hipGetDevice(&dev);
	if (e != hipSuccess) return e;
	if (dev < 0 || dev >= MAX_DEVICES) return hipErrorInvalidDevice;
	void* ptr = nullptr;
	{
		std::lock_guard<std::mutex> g(g_scratchLock);
		if (!g_scratchPtr[dev])
		{	/* the host-side name is not a device address: ask the runtime for this device's copy */
			e = hipGetSymbolAddress(&g_scratchPtr[dev], HIP_SYMBOL(scratch));
			if (e != hipSuccess)
			{
				g_scratchPtr[dev] = nullptr;
				return e;
			}
		}
		ptr = g_scratchPtr[dev];
	}
	return dn_conv3x3_dw(x, Cin, dy, Cout, H, W, dw, db, ptr, s);
}

/* ---- elementwise and layout kernels ------------------------------------ */
/* 16-byte versions of the elementwise kernels below: a thread handles 8
 * channels of one pixel (C a multiple of 8, 16-byte aligned tensors), a grid
 * row (blockIdx.y) one image row, so no per-element index division; the
 * per-element arithmetic is the scalar kernels' (identical results). */
typedef __attribute__((ext_vector_type(8))) unsigned short u16x8;

__device__ __forceinline__ u16x8 ld8(const bf16_t* p) { return *(const u16x8*)p; }
__device__ __forceinline__ void st8(bf16_t* p, u16x8 v) { *(u16x8*)p = v; }

__global__ void k_relu_bwd8(const bf16_t* __restrict__ y, const bf16_t* __restrict__ dy, bf16_t* __restrict__ out, size_t n8)
{
	for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n8; i += (size_t)gridDim.x * blockDim.x)
	{
		const u16x8 a = ld8(y + 8 * i), g = ld8(dy + 8 * i);
		u16x8 o;
#pragma unroll
		for (int k = 0; k < 8; k++) o[k] = bf2f(a[k]) > 0.0f ? g[k] : (unsigned short)0;
		st8(out + 8 * i, o);
	}
}

/* pixel row y of a W-wide image: groups j = px * G + c8 of 8 channels (G = C / 8 = 1 << gs) */
__global__ void k_pool2_bwd8(const bf16_t* __restrict__ dy, int gs, int W, bf16_t* __restrict__ dx, int acc)
{
	const int j = blockIdx.x * blockDim.x + threadIdx.x, G = 1 << gs, py = blockIdx.y;
	if (j >= W * G) return;
	const int px = j >> gs, c8 = j & (G - 1);
	const u16x8 d = ld8(dy + ((size_t)(py >> 1) * (W >> 1) + (px >> 1)) * (8 * G) + 8 * c8);
	bf16_t* o = dx + ((size_t)py * W + px) * (8 * G) + 8 * c8;
	const u16x8 prev = acc ? ld8(o) : u16x8{};
	u16x8 r;
#pragma unroll
	for (int k = 0; k < 8; k++)
	{
		float v = bf2f(d[k]) * 0.25f;
		if (acc) v += bf2f(prev[k]);
		r[k] = f2bf(v);
	}
	st8(o, r);
}

__global__ void k_pool2_8(const bf16_t* __restrict__ x, int gs, int W, bf16_t* __restrict__ y)
{
	const int w2 = W >> 1, j = blockIdx.x * blockDim.x + threadIdx.x, G = 1 << gs, py = blockIdx.y;
	if (j >= w2 * G) return;
	const int px = j >> gs, c8 = j & (G - 1), C = 8 * G;
	const bf16_t* b = x + ((size_t)(2 * py) * W + 2 * px) * C + 8 * c8;
	const u16x8 a = ld8(b), c = ld8(b + C), d = ld8(b + (size_t)W * C), e = ld8(b + (size_t)W * C + C);
	u16x8 r;
#pragma unroll
	for (int k = 0; k < 8; k++) r[k] = f2bf(((bf2f(a[k]) + bf2f(c[k])) + (bf2f(d[k]) + bf2f(e[k]))) * 0.25f);
	st8(y + ((size_t)py * w2 + px) * C + 8 * c8, r);
}

/* cat row: groups j = px * (G0 + G1) + g; the first G0 from the coarse pixel */
__global__ void k_upcat8(const bf16_t* __restrict__ coarse, int G0, const bf16_t* __restrict__ skip, int G1, int W, bf16_t* __restrict__ out)
{
	const int G = G0 + G1, j = blockIdx.x * blockDim.x + threadIdx.x, py = blockIdx.y;
	if (j >= W * G) return;
	const int px = j / G, g = j - px * G;
	const u16x8 v = g < G0 ? ld8(coarse + ((size_t)(py >> 1) * (W >> 1) + (px >> 1)) * (8 * G0) + 8 * g)
		: ld8(skip + ((size_t)py * W + px) * (8 * G1) + 8 * (g - G0));
	st8(out + ((size_t)py * W + px) * (8 * G) + 8 * g, v);
}

__global__ void k_upcat_bwd_coarse8(const bf16_t* __restrict__ dcat, int G0, int G1, int W, bf16_t* __restrict__ dcoarse)
{
	const int G = G0 + G1, w2 = W >> 1, j = blockIdx.x * blockDim.x + threadIdx.x, py = blockIdx.y;
	if (j >= w2 * G0) return;
	const int px = j / G0, g = j - px * G0, C = 8 * G;
	const bf16_t* b = dcat + ((size_t)(2 * py) * W + 2 * px) * C + 8 * g;
	const u16x8 a = ld8(b), c = ld8(b + C), d = ld8(b + (size_t)W * C), e = ld8(b + (size_t)W * C + C);
	u16x8 r;
#pragma unroll
	for (int k = 0; k < 8; k++) r[k] = f2bf((bf2f(a[k]) + bf2f(c[k])) + (bf2f(d[k]) + bf2f(e[k])));
	st8(dcoarse + ((size_t)py * w2 + px) * (8 * G0) + 8 * g, r);
}

__global__ void k_upcat_bwd_skip8(const bf16_t* __restrict__ dcat, int G0, int G1, int W, bf16_t* __restrict__ dskip)
{
	const int G = G0 + G1, j = blockIdx.x * blockDim.x + threadIdx.x, py = blockIdx.y;
	if (j >= W * G1) return;
	const int px = j / G1, g = j - px * G1;
	st8(dskip + ((size_t)py * W + px) * (8 * G1) + 8 * g, ld8(dcat + ((size_t)py * W + px) * (8 * G) + 8 * (G0 + g)));
}

static bool aligned16(const void* p) { return ((uintptr_t)p & 15) == 0; }
static int log2_exact(int v) { int k = 0; while ((1 << k) < v) k++; return (1 << k) == v ? k : -1; }
static dim3 row_grid(int perRow, int rows) { return dim3((unsigned)((perRow + 255) / 256), (unsigned)rows); }

__global__ void k_relu_bwd(const bf16_t* __restrict__ y, const bf16_t* __restrict__ dy, bf16_t* __restrict__ out, size_t n)
{
	for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
		out[i] = bf2f(y[i]) > 0.0f ? dy[i] : (bf16_t)0;
}

static bool aligned16_3(const void* a, const void* b, const void* c) { return (((uintptr_t)a | (uintptr_t)b | (uintptr_t)c) & 15) == 0; }

hipError_t dn_relu_bwd(const bf16_t* y, const bf16_t* dy, bf16_t* dyPre, size_t n, hipStream_t s)
{
	if (!n) return hipSuccess;
	if (n % 8 == 0 && aligned16_3(y, dy, dyPre))
	{
		const size_t n8 = n / 8;
		const int blocks = (int)((n8 + 255) / 256 < 65536 ? (n8 + 255) / 256 : 65536);
		k_relu_bwd8<<<blocks, 256, 0, s>>>(y, dy, dyPre, n8);
		return hipGetLastError();
	}
	const int blocks = (int)((n + 255) / 256 < 65536 ? (n + 255) / 256 : 65536);
	k_relu_bwd<<<blocks, 256, 0, s>>>(y, dy, dyPre, n);
	return hipGetLastError();
}

__global__ void k_pool2(const bf16_t* __restrict__ x, int C, int H, int W, bf16_t* __restrict__ y)
{
	const int h2 = H / 2, w2 = W / 2;
	const size_t n = (size_t)h2 * w2 * C;
	for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
	{
		const int c = (int)(i % C);
		const size_t p = i / C;
		const int px = (int)(p % w2), py = (int)(p / w2);
		const size_t b = ((size_t)(2 * py) * W + 2 * px) * C + c;
		const float v = ((bf2f(x[b]) + bf2f(x[b + C])) + (bf2f(x[b + (size_t)W * C]) + bf2f(x[b + (size_t)W * C + C]))) * 0.25f;
		y[i] = f2bf(v);
	}
}

__global__ void k_pool2_bwd(const bf16_t* __restrict__ dy, int C, int H, int W, bf16_t* __restrict__ dx, int acc)
{
	const size_t n = (size_t)H * W * C;
	for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
	{
		const int c = (int)(i % C);
		const size_t p = i / C;
		const int px = (int)(p % W), py = (int)(p / W);
		float v = bf2f(dy[((size_t)(py / 2) * (W / 2) + px / 2) * C + c]) * 0.25f;
		if (acc) v += bf2f(dx[i]);
		dx[i] = f2bf(v);
	}
}

__global__ void k_upcat(const bf16_t* __restrict__ coarse, int C0, const bf16_t* __restrict__ skip, int C1, int H, int W, bf16_t* __restrict__ out)
{
	const int C = C0 + C1;
	const size_t n = (size_t)H * W * C;
	for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
	{
		const int c = (int)(i % C);
		const size_t p = i / C;
		const int px = (int)(p % W), py = (int)(p / W);
		out[i] = c < C0 ? coarse[((size_t)(py / 2) * (W / 2) + px / 2) * C0 + c] : skip[p * C1 + (c - C0)];
	}
}

__global__ void k_upcat_bwd_coarse(const bf16_t* __restrict__ dcat, int C0, int C1, int H, int W, bf16_t* __restrict__ dcoarse)
{
	const int C = C0 + C1, h2 = H / 2, w2 = W / 2;
	const size_t n = (size_t)h2 * w2 * C0;
	for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
	{
		const int c = (int)(i % C0);
		const size_t p = i / C0;
		const int px = (int)(p % w2), py = (int)(p / w2);
		const size_t b = ((size_t)(2 * py) * W + 2 * px) * C + c;
		dcoarse[i] = f2bf((bf2f(dcat[b]) + bf2f(dcat[b + C])) + (bf2f(dcat[b + (size_t)W * C]) + bf2f(dcat[b + (size_t)W * C + C])));
	}
}

__global__ void k_upcat_bwd_skip(const bf16_t* __restrict__ dcat, int C0, int C1, size_t pixels, bf16_t* __restrict__ dskip)
{
	const int C = C0 + C1;
	const size_t n = pixels * C1;
	for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
		dskip[i] = dcat[(i / C1) * C + C0 + (i % C1)];
}

static int grid_for(size_t n) { const size_t b = (n + 255) / 256; return (int)(b < 65536 ? (b ? b : 1) : 65536); }

hipError_t dn_pool2(const bf16_t* x, int C, int H, int W, bf16_t* y, hipStream_t s)
{
	if (H % 2 || W % 2) return hipErrorInvalidValue;
	if (!H || !W || !C) return hipSuccess;
	const int gs = C % 8 ? -1 : log2_exact(C / 8);
	if (gs >= 0 && aligned16(x) && aligned16(y) && H / 2 <= 65535)
		k_pool2_8<<<row_grid((W / 2) << gs, H / 2), 256, 0, s>>>(x, gs, W, y);
	else k_pool2<<<grid_for((size_t)H * W * C / 4), 256, 0, s>>>(x, C, H, W, y);
	return hipGetLastError();
}

hipError_t dn_pool2_bwd(const bf16_t* dy, int C, int H, int W, bf16_t* dx, int accumulate, hipStream_t s)
{
	if (H % 2 || W % 2) return hipErrorInvalidValue;
	if (!H || !W || !C) return hipSuccess;
	const int gs = C % 8 ? -1 : log2_exact(C / 8);
	if (gs >= 0 && aligned16(dy) && aligned16(dx) && H <= 65535)
		k_pool2_bwd8<<<row_grid(W << gs, H), 256, 0, s>>>(dy, gs, W, dx, accumulate);
	else k_pool2_bwd<<<grid_for((size_t)H * W * C), 256, 0, s>>>(dy, C, H, W, dx, accumulate);
	return hipGetLastError();
}

hipError_t dn_upcat(const bf16_t* coarse, int C0, const bf16_t* skip, int C1, int H, int W, bf16_t* out, hipStream_t s)
{
	if (H % 2 || W % 2) return hipErrorInvalidValue;
	if (!H || !W || !(C0 + C1)) return hipSuccess;
	if (C0 > 0 && C1 > 0 && C0 % 8 == 0 && C1 % 8 == 0 && aligned16_3(coarse, skip, out) && H <= 65535)
		k_upcat8<<<row_grid(W * ((C0 + C1) / 8), H), 256, 0, s>>>(coarse, C0 / 8, skip, C1 / 8, W, out);
	else k_upcat<<<grid_for((size_t)H * W * (C0 + C1)), 256, 0, s>>>(coarse, C0, skip, C1, H, W, out);
	return hipGetLastError();
}

hipError_t dn_upcat_bwd(const bf16_t* dcat, int C0, int C1, int H, int W, bf16_t* dcoarse, bf16_t* dskip, hipStream_t s)
{
	if (H % 2 || W % 2) return hipErrorInvalidValue;
	if (!H || !W || !(C0 + C1)) return hipSuccess;
	if (C0 > 0 && C1 > 0 && C0 % 8 == 0 && C1 % 8 == 0 && aligned16_3(dcat, dcoarse, dskip) && H <= 65535)
	{
		k_upcat_bwd_coarse8<<<row_grid((W / 2) * (C0 / 8), H / 2), 256, 0, s>>>(dcat, C0 / 8, C1 / 8, W, dcoarse);
		k_upcat_bwd_skip8<<<row_grid(W * (C1 / 8), H), 256, 0, s>>>(dcat, C0 / 8, C1 / 8, W, dskip);
		return hipGetLastError();
	}
	k_upcat_bwd_coarse<<<grid_for((size_t)H * W * C0 / 4), 256, 0, s>>>(dcat, C0, C1, H, W, dcoarse);
	k_upcat_bwd_skip<<<grid_for((size_t)H * W * C1), 256, 0, s>>>(dcat, C0, C1, (size_t)H * W, dskip);
	return hipGetLastError();
}

__global__ void k_pack_input(const float4* __restrict__ rad, const float4* __restrict__ alb, const float4* __restrict__ nrm,
	const float* __restrict__ depth, size_t pixels, bf16_t* __restrict__ out)
{
	for (size_t p = blockIdx.x * (size_t)blockDim.x + threadIdx.x; p < pixels; p += (size_t)gridDim.x * blockDim.x)
	{
		const float4 r = rad[p], a = alb[p], n = nrm[p];
		const float d = depth[p];
		auto pk = [](float lo, float hi) { return (uint32_t)f2bf(lo) | (uint32_t)f2bf(hi) << 16; };
		uint4* o = (uint4*)(out + p * 16);
		o[0] = make_uint4(pk(r.x, r.y), pk(r.z, a.x), pk(a.y, a.z), pk(n.x, n.y));
		o[1] = make_uint4(pk(n.z, d), 0, 0, 0);
	}
}

hipError_t dn_pack_input(const float4* radiance, const float4* albedo, const float4* normal, const float* depth, int H, int W, bf16_t* out, hipStream_t s)
{
	k_pack_input<<<grid_for((size_t)H * W), 256, 0, s>>>(radiance, albedo, normal, depth, (size_t)H * W, out);
	return hipGetLastError();
}


__global__ void k_frame_input(const float4* __restrict__ acc, const uint4* __restrict__ feat, float invSpp, int w, int h, int Hp, int Wp,
	bf16_t* __restrict__ out)
{
	const size_t pixels = (size_t)Hp * Wp;
	for (size_t p = blockIdx.x * (size_t)blockDim.x + threadIdx.x; p < pixels; p += (size_t)gridDim.x * blockDim.x)
	{
		const int x = (int)(p % (size_t)Wp), y = (int)(p / (size_t)Wp);
		uint4* o = (uint4*)(out + p * 16);
		if (x >= w || y >= h)
		{
			o[0] = make_uint4(0, 0, 0, 0), o[1] = make_uint4(0, 0, 0, 0);
			continue;
		}
		const size_t q = (size_t)x + (size_t)y * w;
		const float4 d = acc[q], i = acc[q + (size_t)w * h];
		const uint4 f = feat[q];
		const float3 a = feat_albedo(f.x), n = feat_normal(f.y);
		float depth = __uint_as_float(f.z);
		depth = __builtin_isfinite(depth) ? fminf(depth, 1.0e4f) : 1.0e4f;
		const float r[3] = { (d.x + i.x) * invSpp, (d.y + i.y) * invSpp, (d.z + i.z) * invSpp };
		const float al[3] = { a.x, a.y, a.z };
		float c[3];
		for (int k = 0; k < 3; k++) c[k] = log1pf(fmaxf(r[k], 0.0f) / fmaxf(al[k], DN_ALBEDO_FLOOR));
		const float dz = log1pf(depth) * 0.1f;
		auto pk = [](float lo, float hi) { return (uint32_t)f2bf(lo) | (uint32_t)f2bf(hi) << 16; };
		o[0] = make_uint4(pk(c[0], c[1]), pk(c[2], a.x), pk(a.y, a.z), pk(n.x, n.y));
		o[1] = make_uint4(pk(n.z, dz), 0, 0, 0);
	}
}

hipError_t dn_frame_input(const float4* acc32, const uint4* features, float invSpp, int w, int h, int Hp, int Wp, bf16_t* out, hipStream_t s)
{
	if (w <= 0 || h <= 0 || Hp < h || Wp < w) return hipErrorInvalidValue;
	const size_t px = (size_t)Hp * Wp;
	k_frame_input<<<grid_for(px) < 4096 ? grid_for(px) : 4096, 256, 0, s>>>(acc32, features, invSpp, w, h, Hp, Wp, out);
	return hipGetLastError();
}

__global__ void k_frame_output(const float* __restrict__ net, int Wp, const uint4* __restrict__ feat, int w, int h, float4* __restrict__ image)
{
	const size_t pixels = (size_t)w * h;
	for (size_t q = blockIdx.x * (size_t)blockDim.x + threadIdx.x; q < pixels; q += (size_t)gridDim.x * blockDim.x)
	{
		const int x = (int)(q % (size_t)w), y = (int)(q / (size_t)w);
		const float* o = net + ((size_t)x + (size_t)y * Wp) * 16;
		const float3 a = feat_albedo(feat[q].x);
		const float al[3] = { a.x, a.y, a.z };
		float v[3];
		for (int k = 0; k < 3; k++) v[k] = sqrtf(sqrtf(expm1f(fmaxf(o[k], 0.0f)) * fmaxf(al[k], DN_ALBEDO_FLOOR)));
		image[q] = make_float4(v[0], v[1], v[2], 0.0f);
	}
}

hipError_t dn_frame_output(const float* out, int Hp, int Wp, const uint4* features, int w, int h, float4* image, hipStream_t s)
{
	if (w <= 0 || h <= 0 || Hp < h || Wp < w) return hipErrorInvalidValue;
	const size_t px = (size_t)w * h;
	k_frame_output<<<grid_for(px) < 4096 ? grid_for(px) : 4096, 256, 0, s>>>(out, Wp, features, w, h, image);
	return hipGetLastError();
}

__global__ void k_relmse(const float* __restrict__ out, int Cout, const float4* __restrict__ target, size_t pixels, float invN,
	bf16_t* __restrict__ dout, float* __restrict__ loss)
{
	__shared__ float red[256];
	float l = 0.0f;
	for (size_t p = blockIdx.x * (size_t)blockDim.x + threadIdx.x; p < pixels; p += (size_t)gridDim.x * blockDim.x)
	{
		const float4 t4 = target[p];
		const float t[3] = { t4.x, t4.y, t4.z };
		for (int c = 0; c < Cout; c++)
		{
			if (c < 3)
			{
				const float d = out[p * Cout + c] - t[c], den = t[c] * t[c] + 0.01f;
				l += d * d / den;
				dout[p * Cout + c] = f2bf(2.0f * d / den * invN);
			}
			else dout[p * Cout + c] = 0;
		}
	}
	red[threadIdx.x] = l;
	__syncthreads();
	for (int k = 128; k > 0; k >>= 1)
	{
		if ((int)threadIdx.x < k) red[threadIdx.x] += red[threadIdx.x + k];
		__syncthreads();
	}
	if (threadIdx.x == 0) atomicAdd(loss, red[0] * invN);
}

hipError_t dn_relmse(const float* out, int Cout, const float4* target, int H, int W, bf16_t* dout, float* loss, hipStream_t s)
{
	const size_t pixels = (size_t)H * W;
	const int blocks = grid_for(pixels) < 1024 ? grid_for(pixels) : 1024;
	k_relmse<<<blocks, 256, 0, s>>>(out, Cout, target, pixels, 1.0f / (3.0f * (float)pixels), dout, loss);
	return hipGetLastError();
}

__global__ void k_adam(float* __restrict__ p, float* __restrict__ m, float* __restrict__ v, const float* __restrict__ g, size_t n,
	float lr, float b1, float b2, float eps, float c1, float c2, bf16_t* __restrict__ wfwd, bf16_t* __restrict__ wbwd, int Cout, int Cin)
{
	for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
	{
		float w = p[i];
		if (g)
		{
			const float gi = g[i];
			const float mi = b1 * m[i] + (1.0f - b1) * gi, vi = b2 * v[i] + (1.0f - b2) * gi * gi;
			m[i] = mi, v[i] = vi;
			w -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
			p[i] = w;
		}
		if (wfwd) wfwd[i] = f2bf(w);
		if (wbwd)
		{
			const int ci = (int)(i % Cin), co = (int)((i / Cin) % Cout), tap = (int)(i / ((size_t)Cin * Cout));
			wbwd[((size_t)(8 - tap) * Cin + ci) * Cout + co] = f2bf(w);
		}
	}
}

hipError_t dn_adam(float* p, float* m, float* v, const float* g, size_t n, float lr, float b1, float b2, float eps, int t,
	bf16_t* wfwd, bf16_t* wbwd, int Cout, int Cin, hipStream_t s)
{
	if (wbwd && (Cout <= 0 || Cin <= 0 || n != (size_t)9 * Cout * Cin)) return hipErrorInvalidValue;
	const float c1 = 1.0f - powf(b1, (float)t), c2 = 1.0f - powf(b2, (float)t);
	k_adam<<<grid_for(n), 256, 0, s>>>(p, m, v, g, n, lr, b1, b2, eps, c1, c2, wfwd, wbwd, Cout, Cin);
	return hipGetLastError();
}

} // namespace lh
