"""Input path: sensor_msgs/PointCloud2 data blocks -> pcl::PointXYZ float4 rows on the device.

The reference converts with pcl::fromPCLPointCloud2 (Transformer::convertPointCloud2ToPointCloudXYZ,
tsdf_package/src/transformer.cpp:63-67): the FLOAT32 x, y, z fields of every point, row by row, into
PointXYZ {x, y, z, 1}.  The clouds here are built from known coordinates with extra fields, padding inside
points and at row ends, unaligned field offsets and both byte orders, so the expected rows are known by
construction and compared bit for bit (non-finite values included).  A decoded dense scan must integrate to
exactly the map its float4 rows give.
"""
import numpy as np
import pytest

from ros_tsdf_amd import Engine, abi, configs, synth

pytestmark = pytest.mark.gpu


def make_cloud(xyz, width, point_step, offsets, row_pad=0, bigendian=False, extra=True):
    """PointCloud2 data block for points xyz [width*height, 3]: x, y, z as FLOAT32 at offsets, the other
    bytes of each point filled with an intensity-like float and a ring number, rows padded by row_pad bytes."""
    n = xyz.shape[0]
    height = n // width
    assert width * height == n
    row_step = width * point_step + row_pad
    buf = np.full(row_step * height, 0xA5, np.uint8)
    f4 = np.dtype(">f4" if bigendian else "<f4")
    for i in range(n):
        r, c = divmod(i, width)
        base = r * row_step + c * point_step
        if extra and point_step >= max(offsets) + 10:
            # bytes after z: intensity and ring (never read)
            t = max(offsets) + 4
            buf[base + t:base + t + 4] = np.frombuffer(np.array([i * 0.5], f4).tobytes(), np.uint8)
            buf[base + t + 4:base + t + 6] = np.frombuffer(np.uint16(i % 64).tobytes(), np.uint8)
        for k in range(3):
            o = base + offsets[k]
            # a one-element array keeps the byte order (a numpy scalar is always native)
            buf[o:o + 4] = np.frombuffer(np.array([xyz[i, k]], f4).tobytes(), np.uint8)
    return buf, height, row_step


def expected_rows(xyz):
    out = np.ones((xyz.shape[0], 4), np.float32)
    out[:, :3] = xyz.astype(np.float32)
    return out


def decode(eng, buf, width, height, point_step, row_step, offsets, on_device=False, bigendian=False):
    n = width * height
    out = eng.buffer(max(n, 1) * 16)
    src = buf
    if on_device:
        src = eng.buffer(max(buf.nbytes, 16))
        src.upload(buf)
    got = eng.cloud2_to_points(src, width, height, point_step, row_step, offsets, out, is_bigendian=bigendian,
                               data_bytes=buf.nbytes)
    eng.synchronize()
    assert got == n
    rows = np.zeros((n, 4), np.float32)
    if n:
        out.download(rows)
    return rows


@pytest.fixture(scope="module")
def eng():
    e = Engine(heap_blocks=1 << 16, total_buckets=1 << 16, max_batch_points=1 << 17)
    yield e
    e.close()


def _coords(n, seed):
    rng = np.random.default_rng(seed)
    xyz = rng.uniform(-80, 80, size=(n, 3)).astype(np.float32)
    xyz[::17, 0] = np.nan
    xyz[5::23, 2] = np.inf
    xyz[7::29, 1] = -0.0
    return xyz


@pytest.mark.parametrize("on_device", [False, True])
@pytest.mark.parametrize("layout", [
    dict(point_step=32, offsets=(0, 4, 8), row_pad=0, bigendian=False),   # x y z intensity ring pad: aligned
    dict(point_step=13, offsets=(1, 5, 9), row_pad=3, bigendian=False),   # odd step, unaligned fields
    dict(point_step=24, offsets=(8, 0, 16), row_pad=8, bigendian=True),   # permuted fields, big-endian
    dict(point_step=12, offsets=(0, 4, 8), row_pad=0, bigendian=False),   # packed xyz
])
def test_decode_bit_exact(eng, layout, on_device):
    width = 64
    xyz = _coords(width * 37, 3)
    buf, height, row_step = make_cloud(xyz, width, layout["point_step"], layout["offsets"], layout["row_pad"],
                                       layout["bigendian"])
    rows = decode(eng, buf, width, height, layout["point_step"], row_step, layout["offsets"], on_device,
                  layout["bigendian"])
    assert np.array_equal(rows.view(np.uint32), expected_rows(xyz).view(np.uint32))


def test_empty_cloud(eng):
    out = eng.buffer(16)
    assert eng.cloud2_to_points(np.zeros(0, np.uint8), 0, 1, 16, 0, (0, 4, 8), out) == 0
    assert eng.cloud2_to_points(np.zeros(0, np.uint8), 5, 0, 16, 80, (0, 4, 8), out) == 0


def test_layout_errors(eng):
    width = 8
    xyz = _coords(width * 2, 4)
    buf, height, row_step = make_cloud(xyz, width, 16, (0, 4, 8))
    out = eng.buffer(width * height * 16)
    cases = [
        (dict(datatype=8), abi.TSDF_ERR_INVALID_ARG),                      # FLOAT64 fields: no PointXYZ match
        (dict(offsets=(0, 4, 13)), abi.TSDF_ERR_INVALID_ARG),             # z past point_step
        (dict(row_step=width * 16 - 4), abi.TSDF_ERR_INVALID_ARG),        # rows overlap
        (dict(data_bytes=buf.nbytes - 1), abi.TSDF_ERR_INVALID_ARG),      # last point past the data block
        (dict(out=eng.buffer((width * height - 1) * 16)), abi.TSDF_ERR_CAPACITY),
    ]
    for kw, code in cases:
        args = dict(data=buf, width=width, height=height, point_step=16, row_step=row_step, offsets=(0, 4, 8),
                    out=out)
        args.update(kw)
        with pytest.raises(abi.TsdfError) as ei:
            eng.cloud2_to_points(**args)
        assert ei.value.code == code, kw
    # output rows off a 16-byte boundary (in a buffer a row longer, so the rows would still fit)
    import ctypes as C
    lay = abi.Cloud2Layout(width, height, row_step, buf.nbytes, 16, 0, 4, 8, abi.TSDF_FIELD_FLOAT32, 0)
    n = C.c_int64()
    wide = eng.buffer((width * height + 1) * 16)
    assert abi.load().tsdf_cloud2_to_points(eng.h, C.byref(lay), C.c_void_p(buf.ctypes.data), 0,
                                            C.c_void_p(wide.ptr + 8), width * height, C.byref(n)) == \
        abi.TSDF_ERR_INVALID_ARG
    wide.free()
    # the untouched layout still decodes
    assert eng.cloud2_to_points(buf, width, height, 16, row_step, (0, 4, 8), out) == width * height


def test_decoded_scan_integrates_identically(eng):
    """A dense lidar scan sent as a 2048 x 64 PointCloud2 (x, y, z, intensity, ring; 32-byte points) integrates
    to exactly the map of its float4 rows."""
    pts, sensor, _ = synth.make_scan(configs.DENSE, 0)
    n = pts.shape[0]
    width = 2048
    height = n // width
    pts = pts[:width * height]
    point_step = 32
    row_step = width * point_step
    rec = np.zeros((height, width), dtype=[("x", "<f4"), ("y", "<f4"), ("z", "<f4"), ("intensity", "<f4"),
                                          ("ring", "<u2"), ("pad", "V14")])
    rec["x"] = pts[:, 0].reshape(height, width)
    rec["y"] = pts[:, 1].reshape(height, width)
    rec["z"] = pts[:, 2].reshape(height, width)
    rec["intensity"] = 1.5
    rec["ring"] = np.arange(height, dtype=np.uint16)[:, None]
    buf = rec.view(np.uint8).reshape(-1)
    assert buf.nbytes == row_step * height
    decoded = eng.buffer(width * height * 16)
    assert eng.cloud2_to_points(buf, width, height, point_step, row_step, (0, 4, 8), decoded) == width * height
    direct = eng.buffer(width * height * 16)
    rows = np.ones((width * height, 4), np.float32)
    rows[:, :3] = pts[:, :3]
    direct.upload(rows)
    maps = []
    for src in (decoded, direct):
        m = eng.map(configs.DENSE_PARAMS, buckets_log2=14)
        m.integrate(src, width * height, sensor)
        maps.append(m.export())
        m.close()
    (ka, sa, wa), (kb, sb, wb) = maps
    order_a, order_b = np.lexsort(ka.T[::-1]), np.lexsort(kb.T[::-1])
    assert ka.shape[0] > 1000
    assert np.array_equal(ka[order_a], kb[order_b])
    assert np.array_equal(sa[order_a].view(np.uint32), sb[order_b].view(np.uint32))
    assert np.array_equal(wa[order_a].view(np.uint32), wb[order_b].view(np.uint32))
