"""Frames replayed as one hipGraph (BIMP_FLAG_GRAPH) against the same frames issued eagerly:
bit-identical keypoints and response maps for every replay, with eager frames and host
synchronisations between the replays; host and device frames; one and two streams; BF16."""
import numpy as np
import pytest

import oracle_lib as O
from kpmatch import bits

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def B():
    import bimp_amd
    if bimp_amd.device_count() < 1:
        pytest.fail("no GPU visible to the HIP library")
    return bimp_amd


def _maps(B, cs):
    return [cs.get_map(i, B.MAP_DOUBLE).view(np.uint32).copy() for i in range(len(cs.ks))]


@pytest.mark.parametrize("w,h,prec,single", [(640, 480, "fp32", False), (640, 480, "fp32", True),
                                             (1920, 1080, "fp32", False),
                                             (1920, 1080, "bf16", False)])
def test_graph_replays_equal_eager(B, w, h, prec, single):
    opt = B.default_options()
    if prec == "bf16":
        opt.precision = B.PRECISION_BF16
    base = B.FLAG_SINGLE_STREAM if single else 0
    opt.flags = base
    eager = B.CudaStuff(w, h, O.BASE_BANK, opt)
    graph = B.CudaStuff(w, h, O.BASE_BANK, opt)
    graph.set_flags(base | B.FLAG_GRAPH)
    frames = [B.synth_frame(w, h, 60, 70 + i) for i in range(4)]
    for rep in range(2):
        for i, img in enumerate(frames):
            want, n = eager.detect(img, cap=1 << 16)
            got, m = graph.detect(img, cap=1 << 16)  # host frame: uploaded, then the graph
            print(f"{w}x{h} {prec} single={single} rep {rep} frame {i}: {n} keypoints", flush=True)
            assert m == n and np.array_equal(bits(got), bits(want)), (rep, i)
            for a, b in zip(_maps(B, graph), _maps(B, eager)):
                assert np.array_equal(a, b), (rep, i)
    # device frames with a pitch (a new graph: other frame pointer), then host frames again
    pitch = (w + 255) // 256 * 256
    d = B.DeviceBuffer(pitch * h)
    for i, img in enumerate(frames[:2]):
        padded = np.zeros((h, pitch), np.uint8)
        padded[:, :w] = img
        d.upload(padded)
        want, n = eager.detect(img, cap=1 << 16)
        got, m = graph.detect_device(d, pitch, cap=1 << 16)
        assert m == n and np.array_equal(bits(got), bits(want)), i
    got, m = graph.detect(frames[3], cap=1 << 16)
    want, n = eager.detect(frames[3], cap=1 << 16)
    assert m == n and np.array_equal(bits(got), bits(want))


def test_graph_timed_frames_and_flag_switch(B):
    w, h = 640, 480
    cs = B.CudaStuff(w, h, O.BASE_BANK)
    img = B.synth_frame(w, h, 40, 0)
    want, n = cs.detect(img)
    d_img = B.DeviceBuffer(w * h)
    d_img.upload(img)
    cap = B.KP_MAX_POINTS
    d_out = B.DeviceBuffer(cap * B.KEYPOINT_DTYPE.itemsize)
    s = B.Stream()
    cs.set_flags(B.FLAG_GRAPH)
    per = cs.time_detect_frames(d_img, w, d_out, cap, 5, s)
    assert np.all(per > 0)
    got = np.zeros(cap, B.KEYPOINT_DTYPE)
    d_out.download(got)
    assert np.array_equal(bits(got[:min(n, cap)]), bits(want[:min(n, cap)]))
    # switching flags rebuilds the graph (one stream now), results unchanged
    cs.set_flags(B.FLAG_GRAPH | B.FLAG_SINGLE_STREAM)
    got2, m = cs.detect(img)
    assert m == n and np.array_equal(bits(got2), bits(want))
    cs.set_flags(0)
    got3, m3 = cs.detect(img)
    assert m3 == n and np.array_equal(bits(got3), bits(want))
