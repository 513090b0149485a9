"""The colour conversion on the GPU (isp_color_convert_device) against the oracle
and, where oracle/_ref is built, the reference's own CPU converter: every byte
of every plane, padding untouched, over every source and destination layout,
matrix and range, odd edge sizes, padded and unaligned planes (the byte-wise
path), full 4K frames, and the ISP's own output converted for an encoder."""
import itertools

import numpy as np
import pytest

import holoscan_isp as isp
from holoscan_isp import _abi, color
from holoscan_isp.color import Picture

import cases
import color_cases as cc
import gpu_support as gs
import oracle_lib

pytestmark = pytest.mark.gpu

FILL = 0xA5


def on_gpu(torch, src_host, dst_fmt, matrix, range_, pad=0, offset=0, align=1):
    """Converts a host source Picture on the GPU; returns the destination Picture (host copy
    in .buf) and its device buffer. `offset` shifts both device buffers off their 256-byte
    alignment; `align` rounds the destination's pitches up to its multiple."""
    w, h = src_host.width, src_host.height
    src_dev = Picture(src_host.format, w, h, device=True, pad=(src_host.layout[0][1] - src_host.layout[0][2]))
    raw = torch.full((src_dev.nbytes + offset,), 0, dtype=torch.uint8, device="cuda")
    raw[offset:] = torch.from_numpy(src_host.buf)
    sview = src_dev.view()
    sview.planes[0].data = raw.data_ptr() + offset
    dst = Picture(dst_fmt, w, h, pad=pad, fill=FILL, align=align)
    draw = torch.full((dst.nbytes + offset,), FILL, dtype=torch.uint8, device="cuda")
    dview = dst.view()
    base = dst.buf.ctypes.data
    for i in range(color.plane_count(dst_fmt)):
        dview.planes[i].data = draw.data_ptr() + offset + (dview.planes[i].data - base)
    color.convert_device(sview, dview, matrix, range_, torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    dst.buf[:] = draw[offset:].cpu().numpy()
    return dst


def oracle_of(src_host, dst_fmt, matrix, range_, pad=0, align=1):
    want = Picture(dst_fmt, src_host.width, src_host.height, pad=pad, fill=FILL, align=align)
    oracle_lib.color_convert(src_host, want, matrix, range_)
    return want


@pytest.mark.parametrize("dst", cc.DESTINATIONS)
def test_every_pair_matrix_range_and_edge(torch_gpu, dst):
    print(f"color dst {dst}", flush=True)
    for src, matrix, range_ in itertools.product(cc.SOURCES, cc.MATRICES, cc.RANGES):
        for i, (w, h) in enumerate(cc.sizes_for(dst)):
            source = cc.rgb_picture(src, w, h, seed=1000 + 100 * dst + 10 * src + i)
            got = on_gpu(torch_gpu, source, dst, matrix, range_)
            want = oracle_of(source, dst, matrix, range_)
            assert np.array_equal(got.buf, want.buf), (src, dst, matrix, range_, w, h)


@pytest.mark.parametrize("pad,offset", [(5, 0), (16, 0), (0, 1), (3, 2), (32, 4)])
def test_padded_and_unaligned_planes(torch_gpu, pad, offset):
    for src, dst in itertools.product(cc.SOURCES, cc.DESTINATIONS):
        w, h = (66, 19) if dst not in cc.PACKED_422 else (66, 19)
        source = cc.rgb_picture(src, w, h, seed=7 + src + 31 * dst, pad=pad)
        got = on_gpu(torch_gpu, source, dst, _abi.MATRIX_BT709, _abi.RANGE_LIMITED, pad=pad, offset=offset)
        want = oracle_of(source, dst, _abi.MATRIX_BT709, _abi.RANGE_LIMITED, pad=pad)
        assert np.array_equal(got.buf, want.buf), (src, dst, pad, offset)


@pytest.mark.parametrize("dst", cc.DESTINATIONS)
def test_vector_spans_beside_ragged_edges(torch_gpu, dst):
    """Pitches rounded up to 16 bytes put the interior spans on the vector loads and stores;
    the last span of a row (a width that is not a multiple of 16) and, for 4:2:0, the odd
    last row take the byte-wise path. Every byte must agree across the seam."""
    for src, (w, h) in itertools.product(cc.SOURCES, cc.VECTOR_SIZES):
        if dst in cc.PACKED_422 and w % 2:
            continue
        source = cc.rgb_picture(src, w, h, seed=11 * w + h + src, align=16)
        got = on_gpu(torch_gpu, source, dst, _abi.MATRIX_BT601, _abi.RANGE_FULL, align=16)
        want = oracle_of(source, dst, _abi.MATRIX_BT601, _abi.RANGE_FULL, align=16)
        assert np.array_equal(got.buf, want.buf), (src, dst, w, h)


@pytest.mark.skipif(not oracle_lib.cc_ref_available(), reason="reference converter not built here")
def test_matches_the_reference_converter(torch_gpu):
    for src, dst in itertools.product(cc.SOURCES, cc.DESTINATIONS):
        for matrix, range_ in itertools.product(cc.MATRICES, cc.RANGES):
            source = cc.rgb_picture(src, 42, 14, seed=src * 7 + dst)
            got = on_gpu(torch_gpu, source, dst, matrix, range_)
            theirs = Picture(dst, 42, 14, fill=FILL)
            oracle_lib.ref_color_convert(source, theirs, matrix, range_)
            assert np.array_equal(got.buf, theirs.buf), (src, dst, matrix, range_)


@pytest.mark.parametrize("src,dst", [(_abi.PIX_RGB8, _abi.PIX_NV12), (_abi.PIX_RGBA8, _abi.PIX_I420),
                                     (_abi.PIX_BGRA8, _abi.PIX_YUYV), (_abi.PIX_RGB8, _abi.PIX_MONO8)])
def test_full_4k_frames(torch_gpu, src, dst):
    print(f"color 4k {src} -> {dst}", flush=True)
    source = cc.rgb_picture(src, 3840, 2160, seed=src + dst)
    got = on_gpu(torch_gpu, source, dst, _abi.MATRIX_BT709, _abi.RANGE_LIMITED)
    want = oracle_of(source, dst, _abi.MATRIX_BT709, _abi.RANGE_LIMITED)
    assert np.array_equal(got.buf, want.buf)


def test_isp_output_converted_for_an_encoder(torch_gpu):
    """The ISP's RGBA8 frame, still in HBM, converted to NV12 on the same stream: the same
    bytes as converting the frame on the host."""
    torch = torch_gpu
    scene, params, out_format = cases.config2(640, 360, seed=5)
    frame = isp.render_scene(scene)
    pipe = isp.IspPipeline(params)
    d_raw = gs.to_device(frame)
    pitch = 4 * frame.width
    rgba = Picture(_abi.PIX_RGBA8, frame.width, frame.height, device=True)
    stream = torch.cuda.current_stream().cuda_stream
    pipe.process(frame.view(d_raw), rgba.buf, pitch, _abi.OUT_RGBA8, None, None, (10_000_000, 1.0), stream)
    nv12 = Picture(_abi.PIX_NV12, frame.width, frame.height, device=True, fill=FILL)
    color.convert_device(rgba, nv12, _abi.MATRIX_BT709, _abi.RANGE_LIMITED, stream)
    torch.cuda.synchronize()
    host_rgba = Picture(_abi.PIX_RGBA8, frame.width, frame.height)
    host_rgba.buf[:] = rgba.buf.cpu().numpy()
    want = oracle_of(host_rgba, _abi.PIX_NV12, _abi.MATRIX_BT709, _abi.RANGE_LIMITED)
    assert np.array_equal(nv12.buf.cpu().numpy(), want.buf)
