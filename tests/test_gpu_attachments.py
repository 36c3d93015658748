"""restir_gbuffer_from_attachments: the reference's G-buffer attachments (gBufferPass.cpp:75-108:
RGBA8 sRGB albedo, RGB16 snorm normal or its fallbacks, RG16 unorm material, RGB32 float world
position, D32 or D24 depth) decoded into the planes, against a numpy decode of the same texels by
the Vulkan sampling rules (bit for bit), with padded rows; and a frame of the pipeline over
attachments encoded from a primary-ray G-buffer equal to the oracle on the decoded planes."""
import numpy as np
import pytest

from restir_amd import abi, scenes
from restir_amd.pipeline import Renderer
from rt_helpers import floats_equal_nan

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch():
    import torch as t
    assert t.cuda.is_available(), "GPU tests need a HIP device"
    return t


def srgb_table():
    v = np.arange(256) / 255.0
    return np.where(v <= 0.04045, v / 12.92, ((v + 0.055) / 1.055) ** 2.4).astype(np.float32)


def decode(fmt, raw, W, H, row):
    """numpy decode of packed texels (raw: uint8 [H * row]) into [H*W, 4] float32."""
    b = raw.reshape(H, row)
    def words(dtype, n):
        size = np.dtype(dtype).itemsize * n
        return b[:, :W * size].copy().view(dtype).reshape(H * W, n)
    out = np.zeros((H * W, 4), np.float32)
    if fmt in ("RGBA8_SRGB", "RGBA8_UNORM"):
        c = words(np.uint8, 4)
        if fmt == "RGBA8_SRGB":
            out[:, :3] = srgb_table()[c[:, :3]]
        else:
            out[:, :3] = c[:, :3].astype(np.float32) / np.float32(255)
        out[:, 3] = c[:, 3].astype(np.float32) / np.float32(255)
    elif fmt in ("RGB16_SNORM", "RGBA16_SNORM"):
        c = words(np.int16, 3 if fmt == "RGB16_SNORM" else 4)[:, :3]
        out[:, :3] = np.maximum(c.astype(np.float32) / np.float32(32767), np.float32(-1))
    elif fmt in ("RGB16_SFLOAT", "RGBA16_SFLOAT"):
        out[:, :3] = words(np.float16, 3 if fmt == "RGB16_SFLOAT" else 4)[:, :3].astype(np.float32)
    elif fmt in ("RGB32_SFLOAT", "RGBA32_SFLOAT"):
        c = words(np.float32, 3 if fmt == "RGB32_SFLOAT" else 4)
        out[:, :c.shape[1]] = c
    elif fmt == "RG16_UNORM":
        out[:, :2] = words(np.uint16, 2).astype(np.float32) / np.float32(65535)
    elif fmt == "D32_SFLOAT":
        out[:, 0] = words(np.float32, 1)[:, 0]
    elif fmt == "D24_UNORM":
        out[:, 0] = (words(np.uint32, 1)[:, 0] & 0xFFFFFF).astype(np.float32) / np.float32(16777215)
    return out


TEXEL_BYTES = dict(RGBA8_SRGB=4, RGBA8_UNORM=4, RGB16_SNORM=6, RGBA16_SNORM=8, RGB16_SFLOAT=6, RGBA16_SFLOAT=8,
                   RGB32_SFLOAT=12, RGBA32_SFLOAT=16, RG16_UNORM=4, D32_SFLOAT=4, D24_UNORM=4)


def random_texels(rng, fmt, W, H, pad):
    tb = TEXEL_BYTES[fmt]
    row = W * tb + pad
    raw = rng.integers(0, 256, size=H * row, dtype=np.uint8)
    b = raw.reshape(H, row)
    if "SFLOAT" in fmt or fmt == "D32_SFLOAT":  # finite values only
        n = tb // (2 if "16" in fmt else 4)
        dt = np.float16 if "16" in fmt else np.float32
        vals = rng.uniform(-4, 4, size=(H, W * n)).astype(dt)
        b[:, :W * tb] = vals.view(np.uint8).reshape(H, W * tb)
    return raw, row


@pytest.mark.parametrize("normal_fmt", ["RGB16_SNORM", "RGBA16_SNORM", "RGB16_SFLOAT", "RGBA16_SFLOAT", "RGB32_SFLOAT"])
@pytest.mark.parametrize("albedo_fmt,depth_fmt", [("RGBA8_SRGB", "D32_SFLOAT"), ("RGBA8_UNORM", "D24_UNORM")])
def test_attachments_decode_like_a_sampler(torch, normal_fmt, albedo_fmt, depth_fmt):
    W, H = 67, 41
    rng = np.random.default_rng(abi.TEXEL[normal_fmt] * 7 + abi.TEXEL[depth_fmt])
    fmts = dict(albedo=albedo_fmt, normal=normal_fmt, material="RG16_UNORM", world_position="RGBA32_SFLOAT",
                depth=depth_fmt)
    host, dev, atts = {}, {}, {}
    for k, f in fmts.items():
        raw, row = random_texels(rng, f, W, H, pad=16 * (1 + abi.TEXEL[f] % 3))
        host[k] = (raw, row)
        dev[k] = torch.from_numpy(raw).cuda()
        atts[k] = (dev[k], f, row)
    ctx = Renderer(scenes.cornell_box(1), 16, 16).ctx
    for nfmt in (abi.NORMAL_F32, abi.NORMAL_F16):
        P = W * H
        g = torch.zeros(abi.GBUFFER_PLANES * P, dtype=torch.float32, device="cuda")
        ctx.gbuffer_from_attachments(atts, W, H, g, normal_format=nfmt)
        torch.cuda.synchronize()
        got = g.cpu().numpy()
        dec = {k: decode(fmts[k], *host[k][:1], W, H, host[k][1]) for k in fmts}
        assert floats_equal_nan(got[0:3 * P], dec["world_position"][:, :3].T.reshape(-1)).all()
        if nfmt == abi.NORMAL_F32:
            assert floats_equal_nan(got[3 * P:6 * P], dec["normal"][:, :3].T.reshape(-1)).all()
        else:
            halves = got[3 * P:6 * P].view(np.float16)[:3 * P]
            np.testing.assert_array_equal(halves, dec["normal"][:, :3].T.reshape(-1).astype(np.float16))
        assert floats_equal_nan(got[6 * P:10 * P], dec["albedo"].T.reshape(-1)).all()
        assert floats_equal_nan(got[10 * P:12 * P], dec["material"][:, :2].T.reshape(-1)).all()
        assert floats_equal_nan(got[12 * P:13 * P], dec["depth"][:, 0]).all()


def test_attachments_reject_bad_formats_and_rows(torch):
    ctx = Renderer(scenes.cornell_box(1), 16, 16).ctx
    W, H = 8, 8
    buf = torch.zeros(16 * W * H, dtype=torch.uint8, device="cuda")
    g = torch.zeros(abi.GBUFFER_PLANES * W * H, dtype=torch.float32, device="cuda")
    ok = dict(albedo=(buf, "RGBA8_SRGB", 4 * W), normal=(buf, "RGB16_SNORM", 6 * W),
              material=(buf, "RG16_UNORM", 4 * W), world_position=(buf, "RGB32_SFLOAT", 12 * W),
              depth=(buf, "D32_SFLOAT", 4 * W))
    ctx.gbuffer_from_attachments(ok, W, H, g)
    for k, bad in (("albedo", (buf, "RG16_UNORM", 4 * W)), ("depth", (buf, "RGB32_SFLOAT", 12 * W)),
                   ("normal", (buf, "RGB16_SNORM", 6 * W - 2))):
        with pytest.raises(abi.RestirError):
            ctx.gbuffer_from_attachments({**ok, k: bad}, W, H, g)


def encode_reference_attachments(planes, W, H):
    """The reference's attachment bytes for a G-buffer: sRGB8 albedo (nearest code of the linear
    value), snorm16 normal, unorm16 material, float32 position and depth."""
    P = W * H
    pl = planes.reshape(13, P)
    tab = srgb_table()
    alb = np.zeros((P, 4), np.uint8)
    for c in range(3):
        alb[:, c] = np.abs(tab[None, :] - pl[6 + c][:, None]).argmin(axis=1)
    alb[:, 3] = np.where(pl[9] > 0.5, 255, 0)
    nrm = np.round(np.clip(pl[3:6].T, -1, 1) * 32767).astype(np.int16)
    mat = np.round(np.clip(pl[10:12].T, 0, 1) * 65535).astype(np.uint16)
    pos = np.ascontiguousarray(pl[0:3].T, np.float32)
    dep = np.ascontiguousarray(pl[12], np.float32)
    return dict(albedo=(alb, "RGBA8_SRGB", 4 * W), normal=(nrm, "RGB16_SNORM", 6 * W),
                material=(mat, "RG16_UNORM", 4 * W), world_position=(pos, "RGB32_SFLOAT", 12 * W),
                depth=(dep, "D32_SFLOAT", 4 * W))


def test_frame_over_reference_attachments_equals_oracle(torch):
    """Attachments encoded from a primary-ray G-buffer of the 1080p scene (at 320x180), decoded
    on the device; Pass A and two Pass C over the decoded planes equal the oracle on them."""
    from test_gpu_parity import run_frame_and_compare
    sc = scenes.config_scene("terrain")
    W, H = 320, 180
    r = Renderer(sc, W, H, mode="unbiased", k=5, spatial_passes=2)
    cams = [scenes.orbit_camera(sc.camera, f) for f in range(2)]
    scratch = torch.zeros(abi.GBUFFER_PLANES * W * H, dtype=torch.float32, device="cuda")
    r.render_gbuffer = lambda cam, i, stream=None: None  # the frame's planes are decoded below
    for cam in cams:
        # the frame's G-buffer arrives as the reference's attachments
        r.ctx.primary_gbuffer(cam, r.dev.d_tri_material, r.dev.d_materials, len(sc.materials), scratch, W, H)
        enc = encode_reference_attachments(scratch.cpu().numpy(), W, H)
        keep = {k: torch.from_numpy(np.ascontiguousarray(v[0]).view(np.uint8).reshape(-1).copy()).cuda()
                for k, v in enc.items()}
        atts = {k: (keep[k], enc[k][1], enc[k][2]) for k in enc}
        r.ctx.gbuffer_from_attachments(atts, W, H, r.gbuf[1 - r.cur])  # the slot begin_frame takes
        torch.cuda.synchronize()
        run_frame_and_compare(r, cam)
