"""GPU: the Tsdf message consumer's marker split (tsdf_msg_markers, the debug node of README.md:80-85).

The split is checked against numpy's boolean partition of the same records, bit for bit and in message order:
on a real message published from dense scans (from host memory and from device memory), on synthetic
messages spanning many 4096-record tiles with every sign pattern at tile edges, on signed zeros, and on an
empty message.
"""
import numpy as np
import pytest

from ros_tsdf_amd import abi, configs, synth

pytestmark = pytest.mark.gpu


def expect(msg):
    pos = msg[:, 0] >= 0
    return msg[pos][:, 2:5], msg[~pos][:, 2:5]


def same(a, b):
    return a.shape == b.shape and np.array_equal(a.view(np.uint32), b.view(np.uint32))


@pytest.fixture(scope="module")
def eng():
    from ros_tsdf_amd import Engine
    e = Engine(heap_blocks=1 << 17, total_buckets=1 << 18, max_batch_points=1 << 17)
    yield e
    e.close()


def test_markers_of_a_published_message(eng):
    m = eng.map(configs.DENSE_PARAMS, buckets_log2=15)
    buf = eng.buffer(16 << 17)
    s = None
    for k, pts, s in synth.scans(configs.DENSE, stop=2):
        buf.upload(pts)
        m.integrate(buf, pts.shape[0], s)
    msg, n = m.publish_msg(s)
    assert n == msg.shape[0] > 1000
    g, r = eng.msg_markers(msg)
    eg, er = expect(msg)
    assert same(g, eg) and same(r, er) and g.shape[0] > 0 and r.shape[0] > 0
    # the same message from device memory
    dmsg = eng.buffer(msg.nbytes)
    dmsg.upload(msg)
    g2, r2 = eng.msg_markers(dmsg)
    assert same(g2, eg) and same(r2, er)
    dmsg.free()
    buf.free()
    m.close()


@pytest.mark.parametrize("n", [1, 4095, 4096, 4097, 3 * 4096 + 17, 1000003])
def test_markers_of_synthetic_messages(eng, n):
    rng = np.random.default_rng(n)
    msg = rng.standard_normal((n, 5)).astype(np.float32)
    # runs of one sign across tile edges
    if n > 8192:
        msg[4000:4200, 0] = np.abs(msg[4000:4200, 0])
        msg[8100:8300, 0] = -np.abs(msg[8100:8300, 0])
    g, r = eng.msg_markers(msg)
    eg, er = expect(msg)
    assert same(g, eg) and same(r, er)


def test_markers_signed_zero_and_empty(eng):
    msg = np.array([[0.0, 1, 1, 2, 3], [-0.0, 1, 4, 5, 6], [-1e-30, 1, 7, 8, 9], [1e-30, 1, 10, 11, 12]], np.float32)
    g, r = eng.msg_markers(msg)
    assert same(g, msg[[0, 1, 3]][:, 2:5]) and same(r, msg[[2]][:, 2:5])
    g, r = eng.msg_markers(np.zeros((0, 5), np.float32))
    assert g.shape == (0, 3) and r.shape == (0, 3)


def test_markers_argument_errors(eng):
    L = abi.load()
    import ctypes as C
    a, b = C.c_int64(), C.c_int64()
    assert L.tsdf_msg_markers(eng.h, None, 5, 0, None, None, C.byref(a), C.byref(b)) == abi.TSDF_ERR_INVALID_ARG
    assert L.tsdf_msg_markers(eng.h, None, -1, 0, None, None, C.byref(a), C.byref(b)) == abi.TSDF_ERR_INVALID_ARG
