"""Video files with the reference's API (vtem/utils/videoio.py:17-124).

The reference reads and writes videos through OpenCV's codecs (DIVX by default).  No
codec library is part of this build, so the container here is uncompressed AVI: a RIFF
'AVI ' file with one video stream of 24-bit bottom-up BGR frames (BI_RGB 'DIB '
chunks), written from the published RIFF/AVI layout (avih, strh, strf
BITMAPINFOHEADER, movi, idx1).  Such files play in common players and are what OpenCV
writes for fourcc 0.  Compressed AVI streams are rejected with an error.

As in the reference, read_video returns the frames as float64 (frames, height, width)
grey levels sum(BGR * [0.114, 0.587, 0.299]) / 255, or the (frames, height, width, 3) BGR
values with color=True; write_video shifts a negative minimum to 0, scales a maximum
above 1 to 1, writes grey frames to all three channels as 255 v and truncates to uint8.
"""

from __future__ import annotations

import struct

import numpy as np

from . import simpleio

GREY_WEIGHTS_BGR = np.array([0.114, 0.587, 0.299])


def _chunk(fid: bytes, payload: bytes) -> bytes:
    pad = b"\0" if len(payload) % 2 else b""
    return fid + struct.pack("<I", len(payload)) + payload + pad


def _list(kind: bytes, payload: bytes) -> bytes:
    return _chunk(b"LIST", kind + payload)


def _frame_bytes(frame: np.ndarray) -> bytes:
    """One (height, width, 3) uint8 BGR frame as bottom-up rows padded to 4 bytes."""
    h, w, _ = frame.shape
    stride = (3 * w + 3) // 4 * 4
    buf = np.zeros((h, stride), np.uint8)
    buf[:, :3 * w] = frame[::-1].reshape(h, 3 * w)
    return buf.tobytes()


def write_avi(frames: np.ndarray, filename, fps: float = 10) -> None:
    """Uncompressed 24-bit AVI of (frames, height, width, 3) uint8 BGR frames."""
    frames = np.asarray(frames)
    if frames.ndim != 4 or frames.shape[3] != 3 or frames.dtype != np.uint8:
        raise ValueError("write_avi takes (frames, height, width, 3) uint8 BGR frames")
    n, h, w, _ = frames.shape
    stride = (3 * w + 3) // 4 * 4
    size = stride * h
    rate, scale = (int(round(fps * 1000)), 1000) if fps != int(fps) else (int(fps), 1)
    usec = int(round(1e6 / fps))
    avih = struct.pack("<14I", usec, size * rate // scale, 0, 0x10, n, 0, 1, size, w, h, 0, 0, 0, 0)
    strh = struct.pack("<4s4sIHHIIIIIIIIhhhh", b"vids", b"DIB ", 0, 0, 0, 0, scale, rate, 0, n, size,
                       0xFFFFFFFF, 0, 0, 0, w, h)
    strf = struct.pack("<IiiHHIIiiII", 40, w, h, 1, 24, 0, size, 0, 0, 0, 0)
    hdrl = _list(b"hdrl", _chunk(b"avih", avih) + _list(b"strl", _chunk(b"strh", strh) + _chunk(b"strf", strf)))
    movi_parts, idx, off = [], [], 4
    for f in range(n):
        c = _chunk(b"00db", _frame_bytes(frames[f]))
        movi_parts.append(c)
        idx.append(struct.pack("<4sIII", b"00db", 0x10, off, size))
        off += len(c)
    movi = _list(b"movi", b"".join(movi_parts))
    body = b"AVI " + hdrl + movi + _chunk(b"idx1", b"".join(idx))
    with open(filename, "wb") as fh:
        fh.write(b"RIFF" + struct.pack("<I", len(body)) + body)


def _walk(data: bytes, start: int, end: int):
    """(id, payload start, payload size, list kind or None) for the chunks in [start, end)."""
    p = start
    while p + 8 <= end:
        fid = data[p:p + 4]
        n = struct.unpack_from("<I", data, p + 4)[0]
        if fid in (b"LIST", b"RIFF"):
            yield fid, p + 12, n - 4, data[p + 8:p + 12]
        else:
            yield fid, p + 8, n, None
        p += 8 + n + (n & 1)


def read_avi(filename):
    """(frames (n, height, width, 3) uint8 BGR, fps) of an uncompressed 24-bit AVI."""
    with open(filename, "rb") as fh:
        data = fh.read()
    if data[:4] != b"RIFF" or data[8:12] != b"AVI ":
        raise ValueError(f"{filename} is not an AVI file")
    w = h = None
    fps = None
    bottom_up = True
    frames = []

    def visit(start, end):
        nonlocal w, h, fps, bottom_up
        for fid, ps, n, kind in _walk(data, start, end):
            if kind is not None:
                visit(ps, ps + n)
            elif fid == b"strh" and data[ps:ps + 4] == b"vids":
                scale, rate = struct.unpack_from("<II", data, ps + 20)
                fps = rate / scale if scale else None
            elif fid == b"strf" and w is None:
                _, w, hh, _, bits, comp = struct.unpack_from("<IiiHHI", data, ps)
                if bits != 24 or comp not in (0, 0x20424944):          # BI_RGB or 'DIB '
                    raise NotImplementedError("compressed AVI streams need a codec library, which this "
                                              "build does not have; only uncompressed 24-bit AVI is read")
                bottom_up = hh > 0
                h = abs(hh)
            elif fid[2:] in (b"db", b"dc") and w is not None and n > 0:
                stride = (3 * w + 3) // 4 * 4
                rows = np.frombuffer(data, np.uint8, stride * h, ps).reshape(h, stride)[:, :3 * w]
                f = rows.reshape(h, w, 3)
                frames.append(f[::-1] if bottom_up else f)

    visit(12, len(data))
    if w is None:
        raise ValueError(f"{filename} has no video stream")
    return np.stack(frames) if frames else np.zeros((0, h, w, 3), np.uint8), fps


def read_video(filename, write_to_file: bool = False, color: bool = False) -> np.ndarray:
    """Frames of a video file as float64: grey (frames, height, width) in [0, 1] by the
    reference's BGR weights, or (frames, height, width, 3) BGR values with color=True
    (vtem/utils/videoio.py:17-66).  write_to_file also stores the array next to the video
    (name with the extension replaced by .h5, as an array file)."""
    bgr, _ = read_avi(filename)
    if bgr.shape[0] == 0:
        raise TypeError(f"Could not read from {filename}.")
    if color:
        vid = bgr.astype(np.float64)
    else:
        vid = np.sum(bgr.astype(np.float64) * GREY_WEIGHTS_BGR, axis=-1) / 255
    if write_to_file:
        simpleio.write_memory_to_file(vid, str(filename).rpartition(".")[0] + ".h5")
    return vid


def write_video(vid_arr, filename, fps: float = 10, h5input: bool = False, fourcc=None) -> None:
    """Write a video (frames, height, width) grey or (frames, height, width, 3) BGR
    (vtem/utils/videoio.py:70-124): a negative minimum is shifted to 0, a maximum above 1
    scaled to 1, grey values go to all three channels as 255 v, and the frames are
    truncated to uint8.  h5input: vid_arr is the name of an array file.  fourcc must be
    None or 0 (uncompressed)."""
    if fourcc not in (None, 0):
        raise NotImplementedError("only uncompressed AVI is written (fourcc None or 0): this build "
                                  "has no video codec library")
    if h5input:
        vid = simpleio.read_file(vid_arr)
    else:
        vid = simpleio._host(vid_arr)
    vid = np.asarray(vid, np.float64)
    if vid.min() < 0:
        vid = vid - vid.min()
    if vid.max() > 1:
        vid = vid / vid.max()
    if vid.ndim == 3:
        vid = np.repeat((255 * vid)[..., None], 3, axis=3)
    write_avi(vid.astype(np.uint8), filename, fps)
