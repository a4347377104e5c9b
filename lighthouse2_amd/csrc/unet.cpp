/* unet.cpp - forward pass of the learned denoiser's U-Net (see unet.h):
 * encoder level l: a = relu(conv(in_l)), b = relu(conv(a)), in_{l+1} = avgpool2(b)
 * (b and its pool from one kernel); bottleneck: a, b; decoder level l = 3..0:
 * c = relu(conv([upsample2(prev), b_l])) read from the two tensors,
 * d = relu(conv(c)); output conv(d_0) in fp32. */
#include "unet.h"
#include "../../include/lh2hip.h"

namespace lh {

static void dnCheck(hipError_t e, const char* what)
{
	if (e == hipErrorOutOfMemory) throw CoreError(LH2HIP_ERR_OOM, std::string(what) + ": out of device memory");
	if (e != hipSuccess) throw CoreError(LH2HIP_ERR_DEVICE, std::string(what) + ": " + hipGetErrorString(e));
}

void UNetForward::LayerShape(int i, int& cin, int& cout)
{
	cout = CH;
	if (i < 2 * LEVELS) cin = (i == 0) ? CIN : CH;
	else if (i < LAYERS - 1) cin = ((i - 2 * LEVELS) & 1) ? CH : 2 * CH;
	else cin = CH, cout = COUT;
}

size_t UNetForward::BlobBytes()
{
	size_t n = 0;
	for (int i = 0; i < LAYERS; i++)
	{
		int ci, co;
		LayerShape(i, ci, co);
		n += (size_t)9 * co * ci * 2 + (size_t)co * 4;
	}
	return n;
}

void UNetForward::SetWeights(const void* blob, size_t bytes, hipStream_t s)
{
	if (!blob || bytes != BlobBytes()) throw CoreError(LH2HIP_ERR_INVALID, "denoiser weights: expected " + std::to_string(BlobBytes()) + " bytes");
	/* one device copy; every layer's bias starts 4-byte aligned (weights are an
	 * even number of bf16 values: 9 Cout Cin with Cout, Cin multiples of 16) */
	size_t o = 0;
	for (int i = 0; i < LAYERS; i++)
	{
		int ci, co;
		LayerShape(i, ci, co);
		wOff[i] = o, o += (size_t)9 * co * ci * 2;
		bOff[i] = o, o += (size_t)co * 4;
	}
	weights.upload(blob, bytes, s);
	dnCheck(hipStreamSynchronize(s), "sync");
	ready = true;
}

bf16_t* UNetForward::Input(int Hp, int Wp)
{
	if (Hp <= 0 || Wp <= 0 || Hp % 16 || Wp % 16) throw CoreError(LH2HIP_ERR_INVALID, "denoiser: image size must be a multiple of 16");
	H = Hp, W = Wp;
	for (int l = 0; l < LEVELS; l++)
	{
		const size_t px = (size_t)(H >> l) * (W >> l);
		in[l].ensure(px * (l == 0 ? CIN : CH) * 2);
		a[l].ensure(px * CH * 2), b[l].ensure(px * CH * 2);
		if (l < LEVELS - 1) c[l].ensure(px * CH * 2), d[l].ensure(px * CH * 2);
	}
	out.ensure((size_t)H * W * COUT * 4);
	return in[0].as<bf16_t>();
}

const float* UNetForward::Forward(hipStream_t s)
{
	if (!ready) throw CoreError(LH2HIP_ERR_STATE, "denoiser: no weights");
	if (!H) throw CoreError(LH2HIP_ERR_STATE, "denoiser: no input");
	dnCheck(dn_conv3x3(in[0].as<bf16_t>(), CIN, Wt(0), Bi(0), a[0].p, CH, H, W, DN_RELU, s), "conv");
	Middle(s);
	dnCheck(dn_conv3x3(d[0].as<bf16_t>(), CH, Wt(LAYERS - 1), Bi(LAYERS - 1), out.p, COUT, H, W, DN_OUT_F32, s), "conv");
	return out.as<float>();
}

void UNetForward::DenoiseFrame(const float4* acc32, const uint4* features, float invSpp, int w, int h, float4* image, hipStream_t s)
{
	if (!ready) throw CoreError(LH2HIP_ERR_STATE, "denoiser: no weights");
	const int Hp = (h + 15) & ~15, Wp = (w + 15) & ~15;
	bf16_t* x = Input(Hp, Wp);
	dnCheck(dn_frame_input(acc32, features, invSpp, w, h, Hp, Wp, x, s), "denoiser input");
	dnCheck(dn_conv3x3(x, CIN, Wt(0), Bi(0), a[0].p, CH, H, W, DN_RELU, s), "conv");
	Middle(s);
	dnCheck(dn_conv3x3_display(d[0].as<bf16_t>(), Wt(LAYERS - 1), Bi(LAYERS - 1), features, w, h, image, Hp, Wp, s), "denoiser output conv");
}

void UNetForward::Middle(hipStream_t s)
{
	for (int l = 0; l < LEVELS; l++)
	{
		const int h = H >> l, w = W >> l;
		if (l > 0) dnCheck(dn_conv3x3(in[l].as<bf16_t>(), CH, Wt(2 * l), Bi(2 * l), a[l].p, CH, h, w, DN_RELU, s), "conv");
		if (l < LEVELS - 1)
			dnCheck(dn_conv3x3_pool(a[l].as<bf16_t>(), CH, Wt(2 * l + 1), Bi(2 * l + 1), b[l].as<bf16_t>(), CH, h, w, in[l + 1].as<bf16_t>(), s), "conv");
		else dnCheck(dn_conv3x3(a[l].as<bf16_t>(), CH, Wt(2 * l + 1), Bi(2 * l + 1), b[l].p, CH, h, w, DN_RELU, s), "conv");
	}
	const bf16_t* prev = b[LEVELS - 1].as<bf16_t>();
	for (int l = LEVELS - 2, i = 2 * LEVELS; l >= 0; l--, i += 2)
	{
		const int h = H >> l, w = W >> l;
		dnCheck(dn_conv3x3_cat(prev, b[l].as<bf16_t>(), Wt(i), Bi(i), c[l].p, CH, h, w, DN_RELU, s), "conv");
		dnCheck(dn_conv3x3(c[l].as<bf16_t>(), CH, Wt(i + 1), Bi(i + 1), d[l].p, CH, h, w, DN_RELU, s), "conv");
		prev = d[l].as<bf16_t>();
	}
}

} // namespace lh
