"""Front end: raw data to cpx8 IF samples (write_buff and its tables,
src/sdr_rcv.c:776-936).

Oracle pins (CPU): the tables against independent sign-magnitude decodings of
the MAX2771-style 2-bit and 3-bit codes and hand-computed known answers; the
AGC table against hand-computed roundings and its monotonicity and clipping;
write_buff against hand-decoded bytes, against the block-phase invariant (two
consecutive blocks equal one block of twice the length) and INT8X2 against
the conjugate of CS8.  The library's table generators are host code and are
compared with the oracle's on the CPU; the GPU unpack (marked gpu) is
compared bit for bit with the oracle for every format, odd lengths, block
indices and real-to-complex channels, through the host entry point and the
device one.  The array combine (sdr_arch_combine, src/sdr_array.c:553-579):
the oracle keeps the reference's weight table and is pinned by identity and
quadrature weights, equal-weight sums and the half-away-from-zero rounding;
the GPU kernel forms the products directly and matches it bit for bit.
"""
import numpy as np
import pytest

FMT_INT8, FMT_INT8X2, FMT_RAW8, FMT_RAW16, FMT_RAW16I, FMT_RAW32, FMT_CS8, FMT_CS16 = range(1, 9)
NCH = {FMT_INT8: 1, FMT_INT8X2: 1, FMT_CS8: 1, FMT_CS16: 1, FMT_RAW8: 2, FMT_RAW16: 4,
       FMT_RAW16I: 8, FMT_RAW32: 8}
BPS = {FMT_INT8: 1, FMT_INT8X2: 2, FMT_CS8: 2, FMT_CS16: 4, FMT_RAW8: 1, FMT_RAW16: 2,
       FMT_RAW16I: 2, FMT_RAW32: 4}


def dec(b):
    """cpx8 byte(s) -> (I, Q) by sign extension of the nibbles."""
    b = np.asarray(b, np.uint8).astype(np.int32)
    I = ((b & 0xF) ^ 8) - 8
    Q = ((b >> 4) ^ 8) - 8
    return I, Q


def sm2(bits2, qsign=False):
    """2-bit sign-magnitude: bit 1 sign, bit 0 magnitude 1 / 3 (Q: inverted sign)."""
    mag = np.where(bits2 & 1, 3, 1)
    neg = (bits2 >> 1) & 1
    if qsign:
        neg = 1 - neg
    return np.where(neg == 1, -mag, mag)


def sm3(nib):
    """3-bit code in a nibble: bit 1 sign, magnitude 2 (2 b0 + b3) + 1."""
    mag = 2 * (2 * (nib & 1) + ((nib >> 3) & 1)) + 1
    return np.where((nib >> 1) & 1, -mag, mag)


# ---------------------------------------------------------------------------
# oracle pins
# ---------------------------------------------------------------------------
def test_oracle_lutc_iq_2bit(orc):
    i = np.arange(256)
    for rfch in range(2):
        I, Q = dec(orc.gen_lutc(FMT_RAW8, rfch, 2, 2))
        sh = 4 if rfch % 2 else 0
        assert np.array_equal(I, sm2((i >> sh) & 3))
        assert np.array_equal(Q, sm2((i >> (sh + 2)) & 3, qsign=True))
    # known answers: 0x1B low nibble I bits 11 -> -3, Q bits 10 -> +1;
    # high nibble I bits 01 -> 3, Q bits 00 -> -1
    assert dec(orc.gen_lutc(FMT_RAW16, 0, 2, 2)[0x1B]) == (-3, 1)
    assert dec(orc.gen_lutc(FMT_RAW16, 1, 2, 2)[0x1B]) == (3, -1)


def test_oracle_lutc_i_sampling(orc):
    i = np.arange(256)
    for rfch in range(2):
        sh = 4 if rfch % 2 else 0
        I, Q = dec(orc.gen_lutc(FMT_RAW8, rfch, 1, 2))
        assert np.array_equal(I, sm2((i >> sh) & 3)) and not Q.any()
        I, Q = dec(orc.gen_lutc(FMT_RAW8, rfch, 1, 3))
        assert np.array_equal(I, sm3((i >> sh) & 0xF)) and not Q.any()
    # known answers of the 3-bit code: nibble 0b1001 -> +7, 0b0010 -> -1
    assert dec(orc.gen_lutc(FMT_RAW8, 0, 1, 3)[0x09])[0] == 7
    assert dec(orc.gen_lutc(FMT_RAW8, 0, 1, 3)[0x02])[0] == -1
    # RAW16I: channel k reads the 2-bit field at bit 2 (k % 4), I only
    for rfch in range(8):
        I, Q = dec(orc.gen_lutc(FMT_RAW16I, rfch, 1, 2))
        assert np.array_equal(I, sm2((i >> (2 * (rfch % 4))) & 3)) and not Q.any()


def test_oracle_lutc_rtoc(orc):
    for rfch in range(2):
        lut = orc.gen_lutc_rtoc(rfch)
        I0, Q0 = dec(lut[:256])
        assert not Q0.any()
        assert np.array_equal(I0, sm3((np.arange(256) >> (4 if rfch else 0)) & 0xF))
        for p in range(4):
            I, Q = dec(lut[p * 256:(p + 1) * 256])
            z = I0 * (-1j) ** p              # fs/4 mixing: e^{-j pi p / 2}
            assert np.array_equal(I, np.round(z.real).astype(int))
            assert np.array_equal(Q, np.round(z.imag).astype(int))


def test_oracle_lut8(orc):
    l = orc.gen_lut8(FMT_INT8, 0.0, 1.0)
    assert [l[i] for i in range(8)] == [0, 1, 2, 3, 4, 5, 6, 7]
    assert l[8] == 7 and l[0x80] == -7 and l[0xFF] == -1 and l[0xF9] == -7
    # scale 2, round half up: 1 -> 1, -1 -> 0, 3 -> 2, -3 -> -1, 4 -> 2
    l = orc.gen_lut8(FMT_INT8, 0.0, 2.0)
    assert (l[1], l[0xFF], l[3], l[0xFD], l[4]) == (1, 0, 2, -1, 2)
    # offset: (i - 0.5) / 1 -> round half up
    l = orc.gen_lut8(FMT_CS8, 0.5, 1.0)
    assert (l[0], l[1], l[2], l[0xFF]) == (0, 1, 2, -1)
    # CS16: monotone over the signed range, clipped to +-7
    l = orc.gen_lut8(FMT_CS16, 3.0, 100.0)
    v = l[np.arange(-32768, 32768) & 0xFFFF]
    assert np.all(np.diff(v.astype(int)) >= 0) and v.min() == -7 and v.max() == 7
    assert l[3] == 0 and l[53] == 1 and l[52] == 0


def _luts(orc, fmt, rtoc, rng):
    nch = NCH[fmt]
    if fmt in (FMT_INT8, FMT_INT8X2, FMT_CS8):
        return [orc.gen_lut8(fmt, float(rng.uniform(-3, 3)), float(rng.uniform(4, 20)))]
    if fmt == FMT_CS16:
        return [orc.gen_lut8(fmt, float(rng.uniform(-50, 50)), float(rng.uniform(100, 900))),
                orc.gen_lut8(fmt, float(rng.uniform(-50, 50)), float(rng.uniform(100, 900)))]
    out = []
    for k in range(nch):
        if rtoc[k] and fmt != FMT_RAW16I:
            out.append(orc.gen_lutc_rtoc(k))
        else:
            out.append(orc.gen_lutc(fmt, k, 1 if (fmt == FMT_RAW16I or k % 3 == 1) else 2,
                                    3 if k % 3 == 1 else 2))
    return out


def test_oracle_write_buff_known(orc):
    # RAW16: bytes (0x1B, 0xE4) -> ch0/1 from 0x1B, ch2/3 from 0xE4 (IQ 2-bit)
    luts = [orc.gen_lutc(FMT_RAW16, k, 2, 2) for k in range(4)]
    o = orc.write_buff(FMT_RAW16, np.array([0x1B, 0xE4], np.uint8), 1, 0, [0] * 4, luts)
    got = [dec(o[k][0]) for k in range(4)]
    # 0xE4 = 1110 0100: low nibble I 00 -> 1, Q 01 -> -3; high I 10 -> -1, Q 11 -> 3
    assert got == [(-3, 1), (3, -1), (1, -3), (-1, 3)]
    # INT8 with fs/4 phases from block ix = 1, N = 3: phases 3, 0, 1
    l8 = orc.gen_lut8(FMT_INT8, 0.0, 1.0)
    o = orc.write_buff(FMT_INT8, np.array([2, 3, 5], np.uint8), 3, 1, [1], [l8])
    assert [dec(x) for x in o[0]] == [(0, 2), (3, 0), (0, -5)]
    # CS16 little-endian words
    l8 = orc.gen_lut8(FMT_CS16, 0.0, 1.0)
    o = orc.write_buff(FMT_CS16, np.array([3, 0, 0xFE, 0xFF], np.uint8), 1, 0, [0], [l8, None])
    assert dec(o[0][0]) == (3, -2)


@pytest.mark.parametrize("fmt", list(NCH))
def test_oracle_write_buff_invariants(orc, fmt):
    rng = np.random.default_rng(fmt)
    rtoc = [1, 0, 1, 1, 0, 1, 0, 1]
    luts = _luts(orc, fmt, rtoc, rng)
    N = 37
    raw = rng.integers(0, 256, 2 * N * BPS[fmt], dtype=np.uint8)
    # blocks 4 and 5 of length N = samples N..3N of one block of length 2N at
    # index 2 (same phase origin 4N)
    a = orc.write_buff(fmt, raw[:N * BPS[fmt]], N, 4, rtoc, luts)
    b = orc.write_buff(fmt, raw[N * BPS[fmt]:], N, 5, rtoc, luts)
    c = orc.write_buff(fmt, raw, 2 * N, 2, rtoc, luts)
    assert np.array_equal(np.concatenate([a, b], axis=1), c)
    if fmt == FMT_INT8X2:
        c8 = orc.write_buff(FMT_CS8, raw, 2 * N, 0, rtoc, luts)
        I, Q = dec(c)
        I8, Q8 = dec(c8)
        assert np.array_equal(I, I8) and np.array_equal(Q, -Q8)


# ---------------------------------------------------------------------------
# the library (host table generators on the CPU; the unpack on the GPU)
# ---------------------------------------------------------------------------
def test_lib_tables_match_oracle(orc):
    from pocketsdr_amd import sdr_func as s
    for fmt in (FMT_RAW8, FMT_RAW16, FMT_RAW16I, FMT_RAW32):
        for rfch in range(NCH[fmt]):
            for IQ, bits in ((1, 2), (1, 3), (2, 2)):
                assert np.array_equal(s.raw_gen_lut(fmt, rfch, IQ, bits), orc.gen_lutc(fmt, rfch, IQ, bits))
            assert np.array_equal(s.raw_gen_lut(fmt, rfch, 1, 3, rtoc=True), orc.gen_lutc_rtoc(rfch))
    for fmt in (FMT_INT8, FMT_INT8X2, FMT_CS8, FMT_CS16):
        for off, scale in ((0.0, 1.0), (1.3, 7.7), (-2.5, 0.0), (40.0, 333.0)):
            assert np.array_equal(s.raw_gen_lut8(fmt, off, scale), orc.gen_lut8(fmt, off, scale))


@pytest.mark.gpu
@pytest.mark.parametrize("fmt", list(NCH))
def test_raw_unpack_bitexact(sdr, orc, fmt):
    rng = np.random.default_rng(100 + fmt)
    for rtoc in ([0] * 8, [1, 0, 1, 1, 0, 1, 0, 1], [1] * 8):
        luts = _luts(orc, fmt, rtoc, rng)
        plan = sdr.RawPlan(fmt, luts, rtoc)
        assert plan.nch == NCH[fmt]
        for N, ix in ((1, 0), (15, 3), (16, 1), (4099, 7), (65536 + 33, 12345)):
            raw = rng.integers(0, 256, N * BPS[fmt], dtype=np.uint8)
            got = plan.unpack(raw, N, ix)
            want = orc.write_buff(fmt, raw, N, ix, rtoc, luts)
            assert np.array_equal(got, want), (fmt, rtoc, N, ix)
        plan.close()


@pytest.mark.gpu
def test_raw_unpack_dev_matches_host(sdr, orc):
    import torch
    rng = np.random.default_rng(7)
    fmt, rtoc = FMT_RAW32, [1, 0, 0, 1, 1, 0, 1, 0]
    luts = _luts(orc, fmt, rtoc, rng)
    plan = sdr.RawPlan(fmt, luts, rtoc)
    N, ix = 1 << 20, 5
    raw = rng.integers(0, 256, N * BPS[fmt], dtype=np.uint8)
    draw = torch.from_numpy(raw).cuda()
    dout = torch.zeros((8, N), dtype=torch.uint8, device="cuda")
    plan.unpack_dev(draw.data_ptr(), N, ix, [dout[k].data_ptr() for k in range(8)])
    assert np.array_equal(dout.cpu().numpy(), orc.write_buff(fmt, raw, N, ix, rtoc, luts))
    assert plan.kernel_ms() > 0
    plan.close()


# ---------------------------------------------------------------------------
# array combine (sdr_arch_combine, src/sdr_array.c:553-579)
# ---------------------------------------------------------------------------
def cpx8(I, Q):
    return np.uint8(((Q & 0xF) << 4) | (I & 0xF))


def test_oracle_arch_combine(orc):
    vals = np.array([cpx8(i, q) for i in range(-8, 8) for q in range(-8, 8)], np.uint8)
    I, Q = dec(vals)
    zero = np.zeros_like(vals)
    # identity on channel 0 (Q8 weight 256), channel 1 weighted 0
    assert np.array_equal(orc.arch_combine(np.stack([vals, zero]), [[256, 0], [0, 0]]), vals)
    # weight j: (I, Q) -> (-Q, I), -(-8) clipped to 7
    oI, oQ = dec(orc.arch_combine(np.stack([vals, zero]), [[0, 256], [0, 0]]))
    assert np.array_equal(oI, np.minimum(-Q, 7)) and np.array_equal(oQ, I)
    # two equal channels at half weight sum back to the input
    assert np.array_equal(orc.arch_combine(np.stack([vals, vals]), [[128, 0], [128, 0]]), vals)
    # half weight on one channel: x / 2 rounded half away from zero (C
    # division after adding +-128): 1 -> 1, -1 -> -1, 3 -> 2, -3 -> -2, 2 -> 1
    x = np.array([cpx8(v, 0) for v in (1, -1, 3, -3, 2, 0)], np.uint8)
    oI, _ = dec(orc.arch_combine(np.stack([x, np.zeros_like(x)]), [[128, 0], [0, 0]]))
    assert oI.tolist() == [1, -1, 2, -2, 1, 0]


@pytest.mark.gpu
def test_arch_combine_bitexact(sdr, orc):
    import torch
    rng = np.random.default_rng(3)
    for nrfch, narch, N in ((2, 1, 1), (4, 3, 999), (8, 8, 4096 + 7), (8, 4, 1 << 18)):
        ins = rng.integers(0, 256, (nrfch, N), dtype=np.uint8)
        w = rng.integers(-400, 400, (narch, nrfch, 2)).astype(np.int16)
        want = np.stack([orc.arch_combine(ins, w[m]) for m in range(narch)])
        assert np.array_equal(sdr.arch_combine(ins, w), want), (nrfch, narch, N)
        if N % 16 == 0:
            din = torch.from_numpy(ins).cuda()
            dout = torch.zeros((narch, N), dtype=torch.uint8, device="cuda")
            ms = sdr.arch_combine_dev([din[j].data_ptr() for j in range(nrfch)], N, w,
                                      [dout[m].data_ptr() for m in range(narch)])
            assert ms > 0 and np.array_equal(dout.cpu().numpy(), want)


# ---------------------------------------------------------------------------
# front end feeding acquisition in device memory
# ---------------------------------------------------------------------------
def encode_raw8_ch0(buff):
    """cpx8 samples with 2-bit components (+-1, +-3) -> RAW8 bytes whose low
    nibble carries them in the front end's I/Q code (I: sign bit 1, magnitude
    bit 0; Q the same with the sign inverted)."""
    I, Q = dec(buff)
    ii = np.select([I == 1, I == 3, I == -1, I == -3], [0, 1, 2, 3])
    qq = np.select([Q == -1, Q == -3, Q == 1, Q == 3], [0, 1, 2, 3])
    return (ii | (qq << 2)).astype(np.uint8)


@pytest.mark.gpu
def test_raw_to_search_on_device(sdr, orc):
    """RAW8 bytes -> sdr_raw_unpack_dev -> sdr_search_plan_new_dev on the
    device buffer: the same results as the plan on the host cpx8 buffer, also
    after new samples are written into the device buffer between runs."""
    import torch
    from pocketsdr_amd import workloads
    w = workloads.acq_l1ca(seed=21, n_prn=6, n_vis=3, t_int=3e-3, max_dop=2000.0,
                           code_fft_fn=lambda c, T, fs, N: orc.gen_code_fft(c, T, 0.0, fs, N, N))
    s = w["sig"]
    buff = np.asarray(w["buff"], np.uint8)
    n = (len(buff) + 15) // 16 * 16
    luts = [orc.gen_lutc(FMT_RAW8, k, 2, 2) for k in range(2)]
    rp = sdr.RawPlan(FMT_RAW8, luts, [0, 0])
    draw = torch.zeros(n, dtype=torch.uint8, device="cuda")
    dif = torch.zeros((2, n), dtype=torch.uint8, device="cuda")
    raw = encode_raw8_ch0(buff)
    draw[:len(raw)] = torch.from_numpy(raw).cuda()
    rp.unpack_dev(draw.data_ptr(), len(raw), 0, [dif[0].data_ptr(), dif[1].data_ptr()])
    assert np.array_equal(dif[0, :len(buff)].cpu().numpy(), buff)
    pd = sdr.SearchPlan([(dif[0].data_ptr(), len(buff))], [s], sdr.SEARCH_FP16 | sdr.SEARCH_GRAPH,
                        dev_buffs=True)
    for data in (buff, np.roll(buff, 777)):
        raw = encode_raw8_ch0(data)
        draw[:len(raw)] = torch.from_numpy(raw).cuda()
        rp.unpack_dev(draw.data_ptr(), len(raw), 0, [dif[0].data_ptr(), dif[1].data_ptr()])
        pd.run()
        ph = sdr.SearchPlan([data], [s], sdr.SEARCH_FP16 | sdr.SEARCH_GRAPH)
        ph.run()
        for q in range(s["code_fft"].shape[0]):
            assert np.array_equal(pd.power(0, q), ph.power(0, q))
        c1, i1, d1 = pd.result(0)
        c2, i2, d2 = ph.result(0)
        assert np.array_equal(i1, i2) and np.array_equal(c1, c2) and np.array_equal(d1, d2)
        ph.close()
    pd.close()
    rp.close()


@pytest.mark.gpu
@pytest.mark.parametrize("fmt,chans", [(FMT_RAW32, [(1, 0), (4, 1)]), (FMT_RAW16, [(2, 0), (3, 1)]),
                                       (FMT_RAW8, [(0, 1)]), (FMT_RAW16I, [(5, 0)])])
def test_search_plan_on_raw_ring(sdr, orc, fmt, chans):
    """Search plans on packed raw rings (sdr_search_plan_new_raw): the mixing
    loaders read each raw byte and convert it with the channel's table, so
    the FP32 and FP16 power surfaces and results are bit-identical to plans on
    the cpx8 rings sdr_raw_unpack_dev makes of the same bytes (channels with
    and without the fs/4 phase tables, ring phase 1, blocks across the wrap),
    and the FP32 surface is within 1e-5 of the oracle's write_buff followed
    by its sdr_search_code."""
    import torch
    from pocketsdr_amd import workloads
    fs, T = 12e6, 1e-3
    N = int(fs * T)
    nring = 3 * N + 1232                     # a multiple of 16 samples
    rng = np.random.default_rng(fmt)
    bps = {FMT_RAW8: 1, FMT_RAW16: 2, FMT_RAW16I: 2, FMT_RAW32: 4}[fmt]
    raw = rng.integers(0, 256, nring * bps).astype(np.uint8)
    ph0 = 1
    nch = {FMT_RAW8: 2, FMT_RAW16: 4, FMT_RAW16I: 8, FMT_RAW32: 8}[fmt]
    rtoc = [0] * nch
    for k, r in chans:
        rtoc[k] = r
    luts = [orc.gen_lutc_rtoc(k) if rtoc[k] else orc.gen_lutc(fmt, k, 2, 2) for k in range(nch)]
    draw = torch.from_numpy(raw).cuda()
    # the cpx8 rings of the same bytes with ring sample 0 at phase 1: three
    # samples written as block 3 of length 3 (phases 1, 2, 3), the rest as
    # block 0 (phases 0, 1, ...)
    ring = np.concatenate([orc.write_buff(fmt, raw[:3 * bps], 3, 3, rtoc, luts),
                           orc.write_buff(fmt, raw[3 * bps:], nring - 3, 0, rtoc, luts)], axis=1)
    dif = torch.from_numpy(np.ascontiguousarray(ring)).cuda()
    cds = [workloads.l1ca_code(p) for p in (3, 7, 11)]
    cf = np.stack([orc.gen_code_fft(c, T, 0.0, fs, N, N) for c in cds]).astype(np.complex64)
    sigs = [dict(buff=i, M=2 * N, N=N, T=T, fs=fs, fi=0.0, code_fft=cf, fds=workloads.dop_bins(T, 0.0, 1500.0),
                 n_blk=3, ix0=N // 2 + 17, blk_step=N + 301, prns=[3, 7, 11]) for i in range(len(chans))]
    streams = [dict(raw=draw.data_ptr(), N=nring, fmt=fmt, rfch=k, rtoc=r, lut=luts[k], ph0=ph0) for k, r in chans]
    bufs = [(dif[k].data_ptr(), nring) for k, _ in chans]
    for flag in (sdr.SEARCH_BATCHED, sdr.SEARCH_FP16 | sdr.SEARCH_GRAPH):
        pr = sdr.SearchPlan(None, sigs, flag, raw_streams=streams)
        pc = sdr.SearchPlan(bufs, sigs, flag, dev_buffs=True)
        pr.run()
        pc.run()
        for i in range(len(chans)):
            for q in range(3):
                assert np.array_equal(pr.power(i, q).view(np.uint32), pc.power(i, q).view(np.uint32)), (flag, i, q)
            for a, b in zip(pr.result(i), pc.result(i)):
                assert np.array_equal(a, b)
            if flag == sdr.SEARCH_BATCHED:
                s = sigs[i]
                k = chans[i][0]
                R = np.zeros((len(s["fds"]), s["M"]), np.float32)
                for blk in range(s["n_blk"]):
                    orc.search_code(cf[0], T, ring[k], s["ix0"] + blk * s["blk_step"], fs, 0.0, s["fds"], R)
                P = pr.power(i, 0)
                assert np.linalg.norm(P - R[:, :N]) / np.linalg.norm(R[:, :N]) < 1e-5
        pr.close()
        pc.close()
