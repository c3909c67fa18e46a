"""Writes tests/golden/scene_primitives_labels.npz from the reference's golden image.

Source: examples/baseline/scene_primitives.png in the blazert repository (the image the reference's
`render_other_primitives` ctest compares against; CMakeLists.txt:73-99). The renderer is
examples/scene_primitives/main.cc: FP64, 1080x720, camera (0,5,20), ray direction
(x/W - 0.5, y/H - 0.5, -1), red if prim_id == 0, green for other hits, black for misses; row y is
written to image row H-1-y.

The fixture stores one label per pixel in render order (index y*W + x): 0 miss, 1 prim 0, 2 other prim.
Only this script reads the reference tree (a data file); the test uses the committed .npz.
"""
import os
import struct
import sys
import zlib

import numpy as np

SRC = "/root/reference/examples/baseline/scene_primitives.png"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "scene_primitives_labels.npz")


def decode_png_rgb8(path):
    data = open(path, "rb").read()
    assert data[:8] == b"\x89PNG\r\n\x1a\n"
    pos, idat, w = 8, b"", None
    while pos < len(data):
        (length,) = struct.unpack(">I", data[pos:pos + 4])
        tag = data[pos + 4:pos + 8]
        body = data[pos + 8:pos + 8 + length]
        pos += 12 + length
        if tag == b"IHDR":
            w, h, depth, ctype, _, _, interlace = struct.unpack(">IIBBBBB", body)
            assert depth == 8 and ctype == 2 and interlace == 0, (depth, ctype, interlace)
        elif tag == b"IDAT":
            idat += body
    raw = zlib.decompress(idat)
    bpp, stride = 3, w * 3
    out = np.zeros((h, stride), np.uint8)
    prev = np.zeros(stride, np.int32)
    p = 0
    for y in range(h):
        ft = raw[p]
        line = np.frombuffer(raw[p + 1:p + 1 + stride], np.uint8).astype(np.int32)
        p += 1 + stride
        cur = np.zeros(stride, np.int32)
        if ft == 0:
            cur = line
        elif ft == 2:
            cur = (line + prev) & 255
        else:
            for i in range(stride):
                a = cur[i - bpp] if i >= bpp else 0
                b = prev[i]
                c = prev[i - bpp] if i >= bpp else 0
                if ft == 1:
                    x = a
                elif ft == 3:
                    x = (a + b) >> 1
                else:
                    pa, pb, pc = abs(b - c), abs(a - c), abs(a + b - 2 * c)
                    x = a if (pa <= pb and pa <= pc) else (b if pb <= pc else c)
                cur[i] = (line[i] + x) & 255
        out[y] = cur
        prev = cur
    return out.reshape(h, w, 3)


def main():
    img = decode_png_rgb8(SRC)
    h, w, _ = img.shape
    red = (img[..., 0] == 255) & (img[..., 1] == 0) & (img[..., 2] == 0)
    green = (img[..., 0] == 0) & (img[..., 1] == 255) & (img[..., 2] == 0)
    black = (img == 0).all(axis=2)
    assert (red | green | black).all(), "unexpected colours"
    labels_img = np.where(red, 1, np.where(green, 2, 0)).astype(np.uint8)
    labels = labels_img[::-1].reshape(-1)  # image row H-1-y holds render row y
    print("width", w, "height", h, "black", int(black.sum()), "green", int(green.sum()), "red", int(red.sum()))
    np.savez_compressed(OUT, labels=labels, width=w, height=h)


if __name__ == "__main__":
    sys.exit(main())
