"""The record front-end on the GPU against its oracle (oracle/frontend.py), through the C ABI.

Unpacking (UnpackNRCInput, NRCRecord.glsl:98-125) runs FP32 transcendental and
normalisation arithmetic, so it is compared with the FP64 oracle per field within
FP32 rounding; the epilogue over wire records and the path suffix sums are compared
bit for bit with their float32 restatements; the screen pass within FP32 rounding of
its FP64 restatement.  A frame driven by wire records (unpack, infer, composite,
unpack the back-filled train records, train) is checked against the same frame on
oracle-unpacked records.
"""
import numpy as np
import pytest

import vknrc_amd as v
import vknrc_amd.nrc as n
from oracle import frontend as fe
from vknrc_amd import workloads as wl

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
DEV = "cuda"


def dev_bytes(a):
    return torch.from_numpy(np.ascontiguousarray(a).view(np.uint8).copy()).to(DEV)


def host_as(t, dtype, count):
    torch.cuda.synchronize()
    return t.cpu().numpy().view(dtype)[:count].copy()


@pytest.fixture(scope="module")
def scene():
    s = wl.make_scene(wl.workload_seed(wl.SCENE), primitives=5000, vertices=4000, instances=6, materials=12,
                      textures=3, texture_size=32)
    return s, n.DeviceScene(torch, s["vertices"], s["vertex_indices"], s["texcoords"], s["texcoord_indices"],
                            s["materials"], s["material_ids"], s["transforms"], s["textures"])


def check_queries(q, ref):
    assert np.allclose(q[:, 0:3], ref[:, 0:3], rtol=0, atol=2e-5)          # positions up to ~7
    assert np.array_equal(q[:, 3:5], ref[:, 3:5].astype(np.float32))      # unorm16 / 65535, exact
    # the normal's sphere encoding: acos is steep near the poles, where FP32 normalisation error shows
    d = np.abs(q[:, 5:7] - ref[:, 5:7])
    d[:, 0] = np.minimum(d[:, 0], 1 - d[:, 0])                             # u wraps at the -x axis
    assert np.quantile(d, 0.999) < 1e-5 and d.max() < 2e-3
    assert np.array_equal(q[:, 7], ref[:, 7].astype(np.float32))           # roughness
    assert np.allclose(q[:, 8:14], ref[:, 8:14], rtol=0, atol=2e-5)        # textured and constant colours


def test_unpack_eval_and_train_records_match_oracle(scene):
    s, ds = scene
    rng = np.random.default_rng(801)
    nrec, count = 100003, 99001
    recs = np.zeros(nrec, n.EVAL_RECORD)
    recs["dst"] = rng.integers(0, 2 ** 32, nrec, dtype=np.uint64).astype(np.uint32)
    recs["packed_input"] = wl.make_packed_inputs(s, nrec, rng)
    out = torch.full((nrec, 14), -7.0, dtype=torch.float32, device=DEV)
    cnt = torch.tensor([count], dtype=torch.int32, device=DEV)
    v.unpack_eval_records(ds, dev_bytes(recs), nrec, out, count=cnt)
    q = out.cpu().numpy()
    ref = fe.unpack(s, recs["packed_input"][:count])
    check_queries(q[:count], ref)
    assert np.all(q[count:] == -7.0)                                       # records past the count untouched
    # train records: the same unpack, and target = bias bit for bit
    tr = np.zeros(5000, n.WIRE_TRAIN_RECORD)
    tr["bias"] = rng.lognormal(0, 1, (5000, 3))
    tr["factor"] = rng.uniform(0, 1, (5000, 3))
    tr["packed_input"] = recs["packed_input"][:5000]
    o2 = torch.zeros((5000, 17), dtype=torch.float32, device=DEV)
    v.unpack_train_records(ds, dev_bytes(tr), 5000, o2)
    t = o2.cpu().numpy()
    assert np.array_equal(t[:, :14], q[:5000])
    assert np.array_equal(t[:, 14:17], tr["bias"])


def test_packed_primitives_unpack_the_same_bits(scene):
    """nrc_scene_pack_primitives: the 64-byte records hold the indexed corners, texcoords and material
    ids bit for bit, and unpacking through them gives the same eval and train records as the gather
    through the index arrays."""
    s, _ = scene
    ds = n.DeviceScene(torch, s["vertices"], s["vertex_indices"], s["texcoords"], s["texcoord_indices"],
                       s["materials"], s["material_ids"], s["transforms"], s["textures"])
    ds.pack_primitives()
    P = len(s["material_ids"])
    packed = ds.primitives.cpu().numpy().view(np.float32).reshape(P, 16)
    vi = np.asarray(s["vertex_indices"], np.int64).reshape(P, 3)
    ti = np.asarray(s["texcoord_indices"], np.int64).reshape(P, 3)
    assert np.array_equal(packed[:, :9].reshape(P, 3, 3), np.asarray(s["vertices"], np.float32)[vi])
    assert np.array_equal(packed[:, 9:15].reshape(P, 3, 2), np.asarray(s["texcoords"], np.float32)[ti])
    assert np.array_equal(packed[:, 15].view(np.uint32), np.asarray(s["material_ids"], np.uint32))
    rng = np.random.default_rng(802)
    nrec = 60001
    recs = np.zeros(nrec, n.EVAL_RECORD)
    recs["packed_input"] = wl.make_packed_inputs(s, nrec, rng)
    drec = dev_bytes(recs)
    outs = []
    for on in (False, True):
        ds.use_packed(on)
        o = torch.zeros((nrec, 14), dtype=torch.float32, device=DEV)
        v.unpack_eval_records(ds, drec, nrec, o)
        outs.append(o.cpu().numpy())
    assert np.array_equal(outs[0].view(np.uint32), outs[1].view(np.uint32))
    tr = np.zeros(7000, n.WIRE_TRAIN_RECORD)
    tr["bias"] = rng.lognormal(0, 1, (7000, 3))
    tr["packed_input"] = recs["packed_input"][:7000]
    dtr, touts = dev_bytes(tr), []
    for on in (False, True):
        ds.use_packed(on)
        o = torch.zeros((7000, 17), dtype=torch.float32, device=DEV)
        v.unpack_train_records(ds, dtr, 7000, o)
        touts.append(o.cpu().numpy())
    assert np.array_equal(touts[0].view(np.uint32), touts[1].view(np.uint32))
    # a misaligned packed table is refused before anything launches
    ds.c.d_primitives = ds.primitives.data_ptr() + 4
    with pytest.raises(n.NRCError):
        v.unpack_eval_records(ds, drec, nrec, torch.zeros((nrec, 14), dtype=torch.float32, device=DEV))


def test_composite_records_bit_exact(scene):
    s, _ = scene
    rng = np.random.default_rng(802)
    W, H, B = 97, 61, 512
    ns, nt = 4000, 900
    pix = rng.choice(W * H, min(ns, W * H), replace=False)
    ns = len(pix)
    rec = np.zeros(ns + nt + 7, n.EVAL_RECORD)
    rec["dst"][:ns] = [v.dst_screen(int(p % W), int(p // W)) for p in pix]
    # disjoint record ranges of 1..4 records in each batch
    codes = []
    for b in range(4):
        j = 0
        while j < B - 4 and len(codes) < nt:
            L = int(rng.integers(1, 5))
            codes.append(v.dst_train(b, j, j + L - 1))
            j += L + int(rng.integers(0, 3))
    rec["dst"][ns:ns + len(codes)] = codes
    rec["dst"][ns + len(codes):] = 0xFFFFFFFF
    total = len(rec)
    rgb = rng.uniform(0, 3, (total, 3)).astype(np.float32)
    bf = rng.uniform(0, 1, (H, W, 4)).astype(np.float32)
    fgb = rng.uniform(0, 1, (H, W, 2)).astype(np.float32)
    batches = np.zeros(4 * B, n.WIRE_TRAIN_RECORD)
    batches["bias"] = rng.lognormal(0, 1, (4 * B, 3))
    batches["factor"] = rng.uniform(0, 1, (4 * B, 3))
    d_bf, d_fgb = torch.from_numpy(bf.copy()).to(DEV), torch.from_numpy(fgb).to(DEV)
    d_b = dev_bytes(batches)
    v.composite_records(torch.from_numpy(rgb).to(DEV), dev_bytes(rec), total, bias_factor_r=d_bf, factor_gb=d_fgb,
                        width=W, height=H, batches=d_b, batch_stride=B)
    fe.composite_records(rgb, rec, total, bf, fgb, W, H, W, batches, B)
    assert np.array_equal(d_bf.cpu().numpy(), bf)
    got = host_as(d_b, n.WIRE_TRAIN_RECORD, 4 * B)
    assert np.array_equal(got["bias"], batches["bias"]) and np.array_equal(got["factor"], batches["factor"])


def test_build_train_records_matches_path_suffix(scene):
    s, _ = scene
    rng = np.random.default_rng(803)
    P = wl.make_traced_paths(s, 3000, rng)
    cap = 16384
    d_b = torch.zeros(4 * cap * 40, dtype=torch.uint8, device=DEV)
    d_c = torch.zeros(4, dtype=torch.int32, device=DEV)
    d_e = torch.zeros(4000 * 20, dtype=torch.uint8, device=DEV)
    d_ec = torch.zeros(1, dtype=torch.int32, device=DEV)
    v.build_train_records(dev_bytes(P), len(P), d_b, d_c, cap, d_e, d_ec, 4000)
    counts = d_c.cpu().numpy()
    for b in range(4):
        assert counts[b] == P["bounce"][P["batch"] == b].sum()
    recs = host_as(d_b, n.WIRE_TRAIN_RECORD, 4 * cap)
    ne = int(d_ec.cpu().numpy()[0])
    assert ne == int((P["light_terminate"] == 0).sum())
    evals = host_as(d_e, n.EVAL_RECORD, ne)
    # every path's records are contiguous in its batch; locate them by the packed input of vertex 0
    key = lambda p: (int(p["primitive_id"]), int(p["flip_bit_instance_id"]), int(p["barycentric_2x16u"]),  # noqa: E731
                     int(p["scattered_dir_2x16u"]))
    where = {}
    for b in range(4):
        for j in range(int(counts[b])):
            where.setdefault(key(recs[b * cap + j]["packed_input"]), []).append((b, j))
    ev_by_dst = {int(e["dst"]): e for e in evals}
    for p in P:
        k = int(p["bounce"])
        bias, factor = fe.path_suffix(p["light"], p["color"], k)
        b, j0 = [x for x in where[key(p["input"][0])] if x[0] == int(p["batch"])][0]
        seg = recs[b * cap + j0:b * cap + j0 + k]
        assert np.array_equal(seg["bias"], bias) and np.array_equal(seg["factor"], factor)
        assert all(key(seg["packed_input"][i]) == key(p["input"][i]) for i in range(k))
        code = v.dst_train(b, j0, j0 + k - 1)
        if p["light_terminate"]:
            assert code not in ev_by_dst
        else:
            assert key(ev_by_dst[code]["packed_input"]) == key(p["input"][k - 1])


def test_build_train_records_stops_at_the_capacity(scene):
    s, _ = scene
    rng = np.random.default_rng(804)
    P = wl.make_traced_paths(s, 40, rng)
    P["batch"] = 1
    P["light_terminate"] = 0
    cap = 50
    d_b = torch.zeros(4 * cap * 40, dtype=torch.uint8, device=DEV)
    d_c = torch.zeros(4, dtype=torch.int32, device=DEV)
    d_e = torch.zeros(64 * 20, dtype=torch.uint8, device=DEV)
    d_ec = torch.zeros(1, dtype=torch.int32, device=DEV)
    v.build_train_records(dev_bytes(P), len(P), d_b, d_c, cap, d_e, d_ec, 64)
    c = d_c.cpu().numpy()
    assert c[1] == P["bounce"].sum() > cap and c[0] == c[2] == c[3] == 0
    ne = int(d_ec.cpu().numpy()[0])
    evals = host_as(d_e, n.EVAL_RECORD, ne)
    # paths whose first slot fit wrote records and one eval record each; the ranges tile 0..cap-1
    spans = sorted(((int(e["dst"]) >> 3) & 0x3FFF, int(e["dst"]) >> 17) for e in evals)
    assert spans[0][0] == 0 and spans[-1][1] == cap - 1
    assert all(a[1] + 1 == b[0] for a, b in zip(spans, spans[1:]))


@pytest.mark.parametrize("accumulate", [False, True])
def test_screen_tonemap_matches_oracle(accumulate):
    rng = np.random.default_rng(805)
    W, H, pitch = 131, 77, 140
    acc = None
    d_acc = torch.zeros((H, pitch, 4), dtype=torch.float32, device=DEV)
    for k in range(4 if accumulate else 1):
        color = rng.lognormal(-0.5, 1.0, (H, pitch, 4)).astype(np.float32)
        out = torch.full((H, pitch, 4), -1.0, dtype=torch.float32, device=DEV)
        v.screen_tonemap(torch.from_numpy(color).to(DEV), W, H, out, accumulate_img=d_acc, accumulate=accumulate,
                         accumulate_count=k, pitch=pitch)
        ref, acc_new = fe.screen(color[:, :W], None if acc is None else acc, accumulate=accumulate, count=k)
        o = out.cpu().numpy()
        assert np.allclose(o[:, :W], ref, rtol=2e-6, atol=1e-6)
        assert np.all(o[:, W:] == -1.0)                                       # the pitch padding is untouched
        if accumulate:
            acc = d_acc.cpu().numpy()[:, :W]
            assert np.allclose(acc, acc_new, rtol=2e-6, atol=0)


def test_wire_record_frame_matches_oracle_unpacked_frame(scene):
    """Eval records -> unpack -> fused inference -> composite; train records back-filled by the
    epilogue -> unpack -> a training step; against the same MLP passes on oracle-unpacked records."""
    s, ds = scene
    rng = np.random.default_rng(806)
    W, H, B = 128, 64, 2048
    ns = W * H
    # train records of each batch and the eval records that back-fill them (one per 2-record range)
    tr = np.zeros(4 * B, n.WIRE_TRAIN_RECORD)
    tr["bias"] = rng.uniform(0, 1, (4 * B, 3))
    tr["factor"] = rng.uniform(0, 0.5, (4 * B, 3))
    tr["packed_input"] = wl.make_packed_inputs(s, 4 * B, rng)
    nt = 4 * B // 2
    rec = np.zeros(ns + nt, n.EVAL_RECORD)
    rec["dst"][:ns] = [v.dst_screen(i % W, i // W) for i in range(ns)]
    rec["dst"][ns:] = [v.dst_train(j // (B // 2), 2 * (j % (B // 2)), 2 * (j % (B // 2)) + 1) for j in range(nt)]
    rec["packed_input"] = wl.make_packed_inputs(s, len(rec), rng)
    total = len(rec)
    st = v.NRCState(v.NRCConfig(seed=807, encoding=v.ENCODING_TRI12))
    q = torch.zeros((total, 14), dtype=torch.float32, device=DEV)
    rgb = torch.zeros((total, 3), dtype=torch.float32, device=DEV)
    d_rec, d_tr = dev_bytes(rec), dev_bytes(tr)
    bf = torch.zeros((H, W, 4), dtype=torch.float32, device=DEV)
    fgb = torch.ones((H, W, 2), dtype=torch.float32, device=DEV)
    v.unpack_eval_records(ds, d_rec, total, q)
    st.infer(q, rgb)
    v.composite_records(rgb, d_rec, total, bias_factor_r=bf, factor_gb=fgb, width=W, height=H, batches=d_tr,
                        batch_stride=B)
    t17 = torch.zeros((4 * B, 17), dtype=torch.float32, device=DEV)
    v.unpack_train_records(ds, d_tr, 4 * B, t17)
    st.train(t17[:B], write_use_weights=True)
    # the same on oracle-unpacked records
    ref_q = torch.from_numpy(fe.unpack(s, rec["packed_input"]).astype(np.float32)).to(DEV)
    st2 = v.NRCState(v.NRCConfig(seed=807, encoding=v.ENCODING_TRI12))
    rgb2 = torch.zeros((total, 3), dtype=torch.float32, device=DEV)
    st2.infer(ref_q, rgb2)
    a, b = rgb.cpu().numpy(), rgb2.cpu().numpy()
    assert np.linalg.norm(a - b) / np.linalg.norm(b) < 2e-3
    # the screen holds factor * pred (bias 0, factors 1 and factor_r 0)
    img = bf.cpu().numpy()
    assert np.array_equal(img[..., 1], a[:ns, 1].reshape(H, W)) and np.all(img[..., 0] == 0)
    # the back-filled targets are bias + factor * pred of the range's eval record
    filled = host_as(d_tr, n.WIRE_TRAIN_RECORD, 4 * B)
    pred = np.repeat(a[ns:], 2, axis=0)
    want = (tr["bias"] + (tr["factor"] * pred).astype(np.float32)).astype(np.float32)
    assert np.array_equal(filled["bias"], want)
    t = t17.cpu().numpy()
    assert np.array_equal(t[:, 14:17], want)
    s1 = st.get_state()
    assert s1["state"]["t"] == 1 and s1["state"]["last_count"] == B and np.isfinite(s1["weight"]).all()
