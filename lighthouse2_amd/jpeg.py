"""Baseline and extended-sequential JPEG (ITU-T T.81, Huffman coded, 8-bit
samples) -> RGBA, for texture files the reference's HostTexture::Load reads
through FreeImage (lib/RenderSystem/host_texture.cpp:200-241), whose JPEG
plugin is IJG libjpeg with its default decoder settings. Those settings fix
every arithmetic step, so this decoder follows the same published methods:

* the accurate integer inverse DCT (Loeffler-Ligtenberg-Moschytz butterfly,
  13 fraction bits for the constants, 2 extra bits between the column and
  the row pass, rounding right shifts, and the output range table that
  clamps overflow up to 255 / down to 0 and wraps beyond it);
* YCbCr -> RGB with 16-bit fixed-point factors 1.402, 0.34414, 0.71414,
  1.772, rounded as libjpeg's conversion tables round them;
* chroma upsampling by the triangle ("fancy") filter: 3/4 of the nearer and
  1/4 of the farther sample with alternating rounding biases, for 2x1 and
  2x2 subsampling.

Progressive images (T.81 Annex G: spectral selection and successive
approximation, DC and AC first and refinement scans with end-of-band runs)
are decoded into the same coefficient arrays and then follow the same path.
Lossless, hierarchical, arithmetic-coded and 12-bit images raise ValueError.
"""
import struct

import numpy as np

# zigzag order: position k of the coded sequence -> row-major index
_ZZ = np.array([0, 1, 8, 16, 9, 2, 3, 10, 17, 24, 32, 25, 18, 11, 4, 5, 12, 19, 26, 33, 40, 48, 41, 34, 27, 20, 13, 6, 7, 14, 21, 28,
                35, 42, 49, 56, 57, 50, 43, 36, 29, 22, 15, 23, 30, 37, 44, 51, 58, 59, 52, 45, 38, 31, 39, 46, 53, 60, 61, 54, 47, 55,
                62, 63], np.int64)

CONST_BITS, PASS1_BITS = 13, 2


def _fix(x):
    return int(x * (1 << CONST_BITS) + 0.5)


F_0_298, F_0_390, F_0_541, F_0_765 = _fix(0.298631336), _fix(0.390180644), _fix(0.541196100), _fix(0.765366865)
F_0_899, F_1_175, F_1_501, F_1_847 = _fix(0.899976223), _fix(1.175875602), _fix(1.501321110), _fix(1.847759065)
F_1_961, F_2_053, F_2_562, F_3_072 = _fix(1.961570560), _fix(2.053119869), _fix(2.562915447), _fix(3.072711026)


def _descale(x, n):
    return (x + (1 << (n - 1))) >> n


def _butterfly(c0, c1, c2, c3, c4, c5, c6, c7, even_shift):
    """One 8-point inverse DCT on int64 arrays; returns the 8 outputs scaled by
    2^(CONST_BITS) (even part shifted by even_shift), before descaling."""
    z1 = (c2 + c6) * F_0_541
    tmp2 = z1 + c6 * (-F_1_847)
    tmp3 = z1 + c2 * F_0_765
    tmp0 = (c0 + c4) << even_shift
    tmp1 = (c0 - c4) << even_shift
    t10, t13, t11, t12 = tmp0 + tmp3, tmp0 - tmp3, tmp1 + tmp2, tmp1 - tmp2
    o0, o1, o2, o3 = c7, c5, c3, c1
    z1, z2, z3, z4 = o0 + o3, o1 + o2, o0 + o2, o1 + o3
    z5 = (z3 + z4) * F_1_175
    o0, o1, o2, o3 = o0 * F_0_298, o1 * F_2_053, o2 * F_3_072, o3 * F_1_501
    z1, z2, z3, z4 = z1 * (-F_0_899), z2 * (-F_2_562), z3 * (-F_1_961), z4 * (-F_0_390)
    z3, z4 = z3 + z5, z4 + z5
    o0, o1, o2, o3 = o0 + z1 + z3, o1 + z2 + z4, o2 + z2 + z3, o3 + z1 + z4
    return (t10 + o3, t11 + o2, t12 + o1, t13 + o0, t13 - o0, t12 - o1, t11 - o2, t10 - o3)


def _range_limit(x):
    """Sample = range table at (x & 1023): x + 128 for 0..127, 255 for
    128..383, 0 for 384..895, x - 896 for 896..1023."""
    m = x & 1023
    return np.where(m < 128, m + 128, np.where(m < 384, 255, np.where(m < 896, 0, m - 896))).astype(np.uint8)


def idct_blocks(coef):
    """coef: (n, 64) dequantized coefficients in row-major order (int64) ->
    (n, 8, 8) uint8 samples."""
    c = np.asarray(coef, np.int64).reshape(-1, 8, 8)
    # pass 1: columns (index [row][col]: the column's 8 entries are c[:, r, col])
    col = _butterfly(*[c[:, r, :] for r in range(8)], CONST_BITS)
    ws = np.stack([_descale(v, CONST_BITS - PASS1_BITS) for v in col], axis=1)          # (n, 8 rows, 8 cols)
    # pass 2: rows
    row = _butterfly(*[ws[:, :, k] for k in range(8)], CONST_BITS)
    out = np.stack([_descale(v, CONST_BITS + PASS1_BITS + 3) for v in row], axis=2)      # (n, 8, 8)
    return _range_limit(out)


def ycc_to_rgb(y, cb, cr):
    """libjpeg's YCbCr -> RGB on uint8 planes."""
    y, cb, cr = (np.asarray(a, np.int64) for a in (y, cb, cr))
    half = 1 << 15
    fx = lambda v: int(v * 65536 + 0.5)
    xb, xr = cb - 128, cr - 128
    r = y + ((fx(1.40200) * xr + half) >> 16)
    b = y + ((fx(1.77200) * xb + half) >> 16)
    g = y + ((-fx(0.34414) * xb + half + -fx(0.71414) * xr) >> 16)
    clamp = lambda v: np.clip(v, 0, 255).astype(np.uint8)
    return clamp(r), clamp(g), clamp(b)


def _upsample_h2(p):
    """Triangle filter, 2x horizontally: out[2i] = (3 p[i] + p[i-1] + 1) >> 2,
    out[2i+1] = (3 p[i] + p[i+1] + 2) >> 2; the outermost outputs copy the
    edge samples."""
    p = p.astype(np.int64)
    h, w = p.shape
    out = np.empty((h, 2 * w), np.int64)
    if w == 1:
        out[:, 0] = out[:, 1] = p[:, 0]
        return out
    left = np.concatenate([p[:, :1], p[:, :-1]], 1)
    right = np.concatenate([p[:, 1:], p[:, -1:]], 1)
    out[:, 0::2] = (3 * p + left + 1) >> 2
    out[:, 1::2] = (3 * p + right + 2) >> 2
    out[:, 0] = p[:, 0]
    out[:, -1] = p[:, -1]
    return out


def _upsample_h2v2(p):
    """Triangle filter, 2x in both directions: column sums 3 p[row] + p[row
    above or below] (edge rows use themselves), then horizontally
    (3 s[i] + s[i-1] + 8) >> 4 and (3 s[i] + s[i+1] + 7) >> 4, the outermost
    columns (4 s + 8) >> 4 and (4 s + 7) >> 4."""
    p = p.astype(np.int64)
    h, w = p.shape
    up = np.concatenate([p[:1], p[:-1]], 0)
    dn = np.concatenate([p[1:], p[-1:]], 0)
    rows = np.empty((2 * h, w), np.int64)
    rows[0::2] = 3 * p + up
    rows[1::2] = 3 * p + dn
    out = np.empty((2 * h, 2 * w), np.int64)
    if w == 1:
        out[:, 0] = (rows[:, 0] * 4 + 8) >> 4
        out[:, 1] = (rows[:, 0] * 4 + 7) >> 4
        return out
    left = np.concatenate([rows[:, :1], rows[:, :-1]], 1)
    right = np.concatenate([rows[:, 1:], rows[:, -1:]], 1)
    out[:, 0::2] = (3 * rows + left + 8) >> 4
    out[:, 1::2] = (3 * rows + right + 7) >> 4
    out[:, 0] = (rows[:, 0] * 4 + 8) >> 4
    out[:, -1] = (rows[:, -1] * 4 + 7) >> 4
    return out


class _Huffman:
    """Canonical Huffman table (T.81 Annex C) with a 16-bit lookahead table:
    entry = (code length << 8) | symbol for every 16-bit window that starts
    with a code."""

    def __init__(self, counts, symbols):
        self.lut = np.zeros(1 << 16, np.int32)
        code, k = 0, 0
        for length in range(1, 17):
            for _ in range(counts[length - 1]):
                lo = code << (16 - length)
                self.lut[lo:lo + (1 << (16 - length))] = (length << 8) | symbols[k]
                code += 1
                k += 1
            code <<= 1
        self.lut = self.lut.tolist()


class _Bits:
    """Bit reader over entropy-coded bytes with FF00 unstuffed; stops at a marker."""

    def __init__(self, data, pos):
        self.d, self.p, self.acc, self.n = data, pos, 0, 0

    def _fill(self):
        d = self.d
        while self.n <= 48:
            b = d[self.p] if self.p < len(d) else 0
            if b == 0xFF:
                nxt = d[self.p + 1] if self.p + 1 < len(d) else 0
                if nxt == 0x00:
                    self.p += 2
                else:                      # a marker: feed zeros, leave the position on it
                    b = 0
            else:
                self.p += 1
            self.acc = (self.acc << 8) | b
            self.n += 8

    def peek16(self):
        if self.n < 16:
            self._fill()
        return (self.acc >> (self.n - 16)) & 0xFFFF

    def skip(self, k):
        self.n -= k
        self.acc &= (1 << self.n) - 1

    def get(self, k):
        if k == 0:
            return 0
        if self.n < k:
            self._fill()
        self.n -= k
        v = self.acc >> self.n
        self.acc &= (1 << self.n) - 1
        return v

    def symbol(self, t):
        e = t.lut[self.peek16()]
        length = e >> 8
        if length == 0:
            raise ValueError("corrupt JPEG data: no Huffman code")
        self.skip(length)
        return e & 255

    def restart(self):
        """Drop the bits of the partial byte and step over an RSTn marker."""
        self.acc, self.n = 0, 0
        d = self.d
        while self.p + 1 < len(d) and not (d[self.p] == 0xFF and 0xD0 <= d[self.p + 1] <= 0xD7):
            self.p += 1
        self.p += 2


def _extend(v, s):
    return v - (1 << s) + 1 if s and v < (1 << (s - 1)) else v


def decode_jpeg(data):
    """JPEG bytes -> (h, w, 4) uint8 RGBA (alpha 255), top row first."""
    data = bytes(data)
    if data[:2] != b"\xff\xd8":
        raise ValueError("not a JPEG file")
    qt, dc_t, ac_t = {}, {}, {}
    frame, restart, adobe_transform = None, 0, None
    comps_coef = None
    pos = 2
    while pos < len(data):
        if data[pos] != 0xFF:
            pos += 1
            continue
        m = data[pos + 1]
        if m in (0xD8, 0x01) or 0xD0 <= m <= 0xD7 or m == 0xFF:
            pos += 1 if m == 0xFF else 2
            continue
        if m == 0xD9:
            break
        seg_len = struct.unpack(">H", data[pos + 2:pos + 4])[0]
        seg = data[pos + 4:pos + 2 + seg_len]
        if m == 0xDB:                                  # DQT
            i = 0
            while i < len(seg):
                pq, tq = seg[i] >> 4, seg[i] & 15
                if pq:
                    vals = np.frombuffer(seg[i + 1:i + 129], ">u2").astype(np.int64)
                    i += 129
                else:
                    vals = np.frombuffer(seg[i + 1:i + 65], np.uint8).astype(np.int64)
                    i += 65
                q = np.zeros(64, np.int64)
                q[_ZZ] = vals
                qt[tq] = q
        elif m == 0xC4:                                # DHT
            i = 0
            while i < len(seg):
                tc, th = seg[i] >> 4, seg[i] & 15
                counts = list(seg[i + 1:i + 17])
                n = sum(counts)
                syms = list(seg[i + 17:i + 17 + n])
                (dc_t if tc == 0 else ac_t)[th] = _Huffman(counts, syms)
                i += 17 + n
        elif m in (0xC0, 0xC1, 0xC2):                  # baseline / extended sequential / progressive, Huffman
            prec, h, w, nc = seg[0], *struct.unpack(">HH", seg[1:5]), seg[5]
            if prec != 8:
                raise ValueError(f"JPEG with {prec}-bit samples: only 8-bit images are read")
            comps = []
            for k in range(nc):
                cid, hv, tq = seg[6 + 3 * k], seg[7 + 3 * k], seg[8 + 3 * k]
                comps.append({"id": cid, "h": hv >> 4, "v": hv & 15, "tq": tq})
            frame = {"h": h, "w": w, "comps": comps, "progressive": m == 0xC2}
            hmax, vmax = max(c["h"] for c in comps), max(c["v"] for c in comps)
            mcux, mcuy = -(-w // (8 * hmax)), -(-h // (8 * vmax))
            for c in comps:
                c["bw"], c["bh"] = mcux * c["h"], mcuy * c["v"]
                c["coef"] = np.zeros((c["bh"], c["bw"], 64), np.int64)
            frame.update(hmax=hmax, vmax=vmax, mcux=mcux, mcuy=mcuy)
        elif m in (0xC3, 0xC5, 0xC6, 0xC7, 0xC9, 0xCA, 0xCB, 0xCD, 0xCE, 0xCF):
            raise ValueError("lossless, hierarchical or arithmetic-coded JPEG images are not read")
        elif m == 0xDD:                                # DRI
            restart = struct.unpack(">H", seg[:2])[0]
        elif m == 0xEE and seg[:5] == b"Adobe":
            adobe_transform = seg[11] if len(seg) > 11 else None
        elif m == 0xDA:                                # SOS
            if frame is None:
                raise ValueError("JPEG scan before the frame header")
            ns = seg[0]
            sc = []
            for k in range(ns):
                cid, t = seg[1 + 2 * k], seg[2 + 2 * k]
                c = next(c for c in frame["comps"] if c["id"] == cid)
                sc.append((c, dc_t.get(t >> 4), ac_t.get(t & 15)))
            if frame["progressive"]:
                ss, se, ah, al = seg[1 + 2 * ns], seg[2 + 2 * ns], seg[3 + 2 * ns] >> 4, seg[3 + 2 * ns] & 15
                if not (ss <= se <= 63) or (ss == 0) != (se == 0) or (ss > 0 and ns != 1):
                    raise ValueError("corrupt JPEG data: bad progressive scan parameters")
                need_dc, need_ac = ss == 0 and ah == 0, ss > 0
            else:
                need_dc = need_ac = True
            if any((need_dc and d is None) or (need_ac and a is None) for _, d, a in sc):
                raise ValueError("corrupt JPEG data: a scan uses an undefined Huffman table")
            if frame["progressive"]:
                pos = _decode_scan_progressive(data, pos + 2 + seg_len, frame, sc, restart, ss, se, ah, al)
            else:
                pos = _decode_scan(data, pos + 2 + seg_len, frame, sc, restart)
            continue
        pos += 2 + seg_len
    if frame is None:
        raise ValueError("JPEG without a frame")
    if len(frame["comps"]) == 4 or adobe_transform == 2:
        raise ValueError("CMYK / YCCK JPEG images are not read")
    h, w = frame["h"], frame["w"]
    planes = []
    for c in frame["comps"]:
        q = qt[c["tq"]]
        blocks = idct_blocks((c["coef"] * q).reshape(-1, 64)).reshape(c["bh"], c["bw"], 8, 8)
        img = blocks.transpose(0, 2, 1, 3).reshape(c["bh"] * 8, c["bw"] * 8)
        fx, fy = frame["hmax"] // c["h"], frame["vmax"] // c["v"]
        cw, ch = -(-w * c["h"] // frame["hmax"]), -(-h * c["v"] // frame["vmax"])
        img = img[:ch, :cw]
        if (fx, fy) == (2, 2):
            img = _upsample_h2v2(img)
        elif (fx, fy) == (2, 1):
            img = _upsample_h2(img)
        elif (fx, fy) != (1, 1):
            img = np.repeat(np.repeat(img, fy, 0), fx, 1)     # other factors: replication
        planes.append(np.asarray(img[:h, :w]))
    rgba = np.full((h, w, 4), 255, np.uint8)
    if len(planes) == 1:
        rgba[..., 0] = rgba[..., 1] = rgba[..., 2] = planes[0].astype(np.uint8)
    elif len(planes) == 3:
        if adobe_transform == 0:                          # stored as RGB
            for k in range(3):
                rgba[..., k] = planes[k].astype(np.uint8)
        else:
            r, g, b = ycc_to_rgb(*planes)
            rgba[..., 0], rgba[..., 1], rgba[..., 2] = r, g, b
    else:
        raise ValueError(f"JPEG with {len(planes)} components is not read")
    return rgba


def _scan_units(frame, sc):
    """The scan's units in coding order (T.81 A.2): lists of (scan component,
    block row, block column). A single-component scan codes the component's
    own blocks (ceil(component width / 8) per row); an interleaved scan codes
    MCUs of h x v blocks per component."""
    if len(sc) == 1:
        c = sc[0][0]
        ncols, nrows = -(-frame["w"] * c["h"] // (8 * frame["hmax"])), -(-frame["h"] * c["v"] // (8 * frame["vmax"]))
        return [[(0, r, q)] for r in range(nrows) for q in range(ncols)]
    units = []
    for my in range(frame["mcuy"]):
        for mx in range(frame["mcux"]):
            u = []
            for k, (c, _, _) in enumerate(sc):
                for by in range(c["v"]):
                    for bx in range(c["h"]):
                        u.append((k, my * c["v"] + by, mx * c["h"] + bx))
            units.append(u)
    return units


def _scan_end(data, p):
    """Steps over what remains of a scan's entropy-coded data, up to the next marker."""
    while p + 1 < len(data) and not (data[p] == 0xFF and data[p + 1] not in (0x00,) and not (0xD0 <= data[p + 1] <= 0xD7)):
        p += 1
    return p


def _decode_scan(data, pos, frame, sc, restart):
    """Huffman-decodes one sequential scan into the components' coefficient
    arrays; returns the position after its entropy-coded data."""
    bits = _Bits(data, pos)
    pred = [0] * len(sc)
    units = _scan_units(frame, sc)
    zz = _ZZ.tolist()
    for n, u in enumerate(units):
        if restart and n and n % restart == 0:
            bits.restart()
            pred = [0] * len(sc)
        for k, r, q in u:
            c, dct, act = sc[k]
            blk = c["coef"][r, q]
            s = bits.symbol(dct)
            pred[k] += _extend(bits.get(s), s)
            blk[0] = pred[k]
            i = 1
            while i < 64:
                rs = bits.symbol(act)
                run, s = rs >> 4, rs & 15
                if s == 0:
                    if run != 15:
                        break
                    i += 16
                    continue
                i += run
                if i > 63:
                    raise ValueError("corrupt JPEG data: coefficient index beyond 63")
                blk[zz[i]] = _extend(bits.get(s), s)
                i += 1
    return _scan_end(data, bits.p)


def _decode_scan_progressive(data, pos, frame, sc, restart, ss, se, ah, al):
    """One progressive scan (T.81 G.1.2) over coefficients ss..se (zigzag
    positions) at bit position al; ah = 0 for a first scan, else the bit
    position of the previous scan of this band (a refinement adds one bit).

    * DC first: the difference code as in a sequential scan, the coefficient
      becomes the running prediction times 2^al.
    * DC refinement: one raw bit per block, ORed in at bit al.
    * AC first: run/size codes as in a sequential scan, values times 2^al;
      a code with size 0 and run r < 15 starts an end-of-band run of
      2^r + (r raw bits) blocks (this one included) with no more coefficients.
    * AC refinement: size is 0 or 1; a new coefficient is +-2^al (one sign
      bit) and lands on the (run + 1)-th still-zero position; every already
      nonzero coefficient passed on the way, and every one after the last code
      of a block in an end-of-band run, reads one correction bit that moves
      it 2^al away from zero unless that bit is already set.
    Returns the position after the scan's entropy-coded data."""
    bits = _Bits(data, pos)
    units = _scan_units(frame, sc)
    zz = _ZZ.tolist()
    pred = [0] * len(sc)
    eobrun = 0
    p1, m1 = 1 << al, -1 << al
    for n, u in enumerate(units):
        if restart and n and n % restart == 0:
            bits.restart()
            pred = [0] * len(sc)
            eobrun = 0
        for k, r, q in u:
            c, dct, act = sc[k]
            blk = c["coef"][r, q]
            if ss == 0:                                   # DC scans (possibly interleaved)
                if ah == 0:
                    s = bits.symbol(dct)
                    pred[k] += _extend(bits.get(s), s)
                    blk[0] = pred[k] * p1
                elif bits.get(1):
                    blk[0] |= p1
                continue
            if ah == 0:                                   # AC first
                if eobrun:
                    eobrun -= 1
                    continue
                i = ss
                while i <= se:
                    rs = bits.symbol(act)
                    run, s = rs >> 4, rs & 15
                    if s:
                        i += run
                        if i > se:
                            raise ValueError("corrupt JPEG data: coefficient index beyond the band")
                        blk[zz[i]] = _extend(bits.get(s), s) * p1
                        i += 1
                    elif run == 15:
                        i += 16
                    else:
                        eobrun = (1 << run) - 1 + bits.get(run)
                        break
                continue
            i = ss                                        # AC refinement
            if eobrun == 0:
                while i <= se:
                    rs = bits.symbol(act)
                    run, s = rs >> 4, rs & 15
                    val = 0
                    if s:
                        if s != 1:
                            raise ValueError("corrupt JPEG data: refinement coefficient of size above 1")
                        val = p1 if bits.get(1) else m1
                    elif run != 15:
                        eobrun = (1 << run) + bits.get(run)
                        break
                    while i <= se:                        # correction bits up to the target zero position
                        z = zz[i]
                        cf = blk[z]
                        if cf:
                            if bits.get(1) and not (cf & p1):
                                blk[z] = cf + p1 if cf >= 0 else cf + m1
                        else:
                            if run == 0:
                                break
                            run -= 1
                        i += 1
                    if val:
                        if i > se:
                            raise ValueError("corrupt JPEG data: coefficient index beyond the band")
                        blk[zz[i]] = val
                    i += 1
            if eobrun:
                while i <= se:                            # correction bits for the rest of the band
                    z = zz[i]
                    cf = blk[z]
                    if cf and bits.get(1) and not (cf & p1):
                        blk[z] = cf + p1 if cf >= 0 else cf + m1
                    i += 1
                eobrun -= 1
    return _scan_end(data, bits.p)
