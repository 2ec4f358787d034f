"""GPU tests of stream ingestion (the reference's `pipe` subcommand,
src/main.rs:1091-1304): frames read from a file descriptor, hashed on the GPU
in batches, written in the reference's output formats.

The records of every format equal the host formatter applied to the CPU
oracle's hashes, qualities, thresholds and orientation PDQFs (x86 AVX2
dihedral semantics, as the reference's x86 binary); the golden image's line
is one of the reference's expected hashes; uneven batches, an empty input and
an input that ends inside a frame behave as specified; the command-line
binary gives the same bytes."""
import os
import subprocess
import sys

import numpy as np
import pytest

from yume_pdq import synthetic

pytestmark = pytest.mark.gpu

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(REPO, "tests", "golden")
PIPE_BIN = os.path.join(REPO, "yume-pdq_amd", "_build", "yume-pdq-pipe")


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def frames():
    golden = np.load(os.path.join(GOLDEN, "aaa_orig_512_u8.npy"))
    imgs = synthetic.images_u8(10, start=300)
    return np.concatenate([golden[None], imgs]).astype(np.uint8)


@pytest.fixture(scope="module")
def oracle_out(oracle, frames):
    """Oracle hashes, qualities, thresholds and orientation PDQFs (AVX2
    dihedral semantics) of every frame."""
    n = frames.shape[0]
    H = np.zeros((n, 8, 32), np.uint8)
    Q = np.zeros((n, 8), np.float32)
    T = np.zeros((n, 8), np.float32)
    P = np.zeros((n, 8, 16, 16), np.float32)
    for i in range(n):
        f = frames[i].astype(np.float32)
        H[i, 0], T[i, 0], Q[i, 0], P[i, 0], _, _ = oracle.hash_get_threshold(f)
        codes, th, q, h, pd = oracle.visit_dihedrals(P[i, 0], T[i, 0], 1)
        H[i, 1:], T[i, 1:], Q[i, 1:], P[i, 1:] = h, th, q, pd
    return H, Q, T, P


def run_pipe(gpu, tmp_path, data, fmt, **kw):
    src = tmp_path / "in.raw"
    dst = tmp_path / "out.txt"
    src.write_bytes(data)
    n = gpu.pipe(str(src), str(dst), fmt=fmt, **kw)
    return n, dst.read_bytes()


FORMATS = ["raw", "q+raw", "hex+lf", "HEX+crlf", "q+hex+lf", "q+HEX", "bin", "q+bin+nul", "diagnostic+lf"]


@pytest.mark.parametrize("fmt", FORMATS)
def test_formats_equal_oracle(gpu, oracle_out, frames, tmp_path, fmt):
    H, Q, T, P = oracle_out
    for dih in (False, True):
        say(f"[pipe] {fmt} dihedrals={dih}")
        n, got = run_pipe(gpu, tmp_path, frames.tobytes(), fmt, dihedrals=dih, batch=4)
        assert n == frames.shape[0]
        want = gpu.format_frames(fmt, H, Q, T, dihedrals=dih, dihedral_pdqf=P)
        assert got == want


def test_golden_line(gpu, frames, tmp_path):
    n, got = run_pipe(gpu, tmp_path, frames[:1].tobytes(), "hex+lf")
    expect = [x.strip() for x in open(os.path.join(GOLDEN, "aaa_orig_expect.txt")) if x.strip()]
    assert got.decode().strip() in expect


@pytest.mark.parametrize("batch", [1, 3, 64])
def test_batches_equal(gpu, frames, tmp_path, batch):
    _, ref = run_pipe(gpu, tmp_path, frames.tobytes(), "q+hex+lf", dihedrals=True, batch=7)
    _, got = run_pipe(gpu, tmp_path, frames.tobytes(), "q+hex+lf", dihedrals=True, batch=batch)
    assert got == ref


def test_empty_input(gpu, tmp_path):
    n, got = run_pipe(gpu, tmp_path, b"", "hex+lf")
    assert n == 0 and got == b""


def test_input_ending_inside_a_frame(gpu, frames, tmp_path):
    """Every whole frame is written, then the call reports the truncated one."""
    data = frames[:3].tobytes() + frames[3].tobytes()[:1000]
    src = tmp_path / "in.raw"
    dst = tmp_path / "out.txt"
    src.write_bytes(data)
    with pytest.raises(gpu.YumePdqError, match="inside a frame"):
        gpu.pipe(str(src), str(dst), fmt="hex+lf", batch=2)
    lines = dst.read_bytes().decode().splitlines()
    _, full = run_pipe(gpu, tmp_path, frames[:3].tobytes(), "hex+lf")
    assert lines == full.decode().splitlines()


@pytest.mark.parametrize("channels", [3, 4])
def test_colour_inputs(gpu, dev, frames, tmp_path, channels):
    """rgb8 / rgba8 frames: the records equal hashing the same pixels with
    the fused colour kernel (itself equal to converting with the reference's
    cvt_*_to_luma8f and hashing)."""
    import torch
    rng = np.random.default_rng(channels)
    px = np.empty(frames.shape + (channels,), np.uint8)
    for c in range(channels):
        px[..., c] = np.clip(frames.astype(np.int16) + rng.integers(-30, 31, size=(frames.shape[0], 1, 1)), 0,
                             255).astype(np.uint8)
    fmt_in = "rgb8" if channels == 3 else "rgba8"
    _, got = run_pipe(gpu, tmp_path, px.tobytes(), "q+hex+lf", dihedrals=True, input_format=fmt_in, batch=5)
    h, q, t = gpu.hash_batch(torch.from_numpy(px).to(dev), dihedral_mode=gpu.DIHEDRAL_AVX2)
    want = gpu.format_frames("q+hex+lf", h.cpu().numpy(), q.cpu().numpy(), t.cpu().numpy(), dihedrals=True)
    assert got == want


def test_command_line(gpu, frames, tmp_path):
    src = tmp_path / "in.raw"
    src.write_bytes(frames.tobytes())
    say("[pipe] command line")
    r = subprocess.run([PIPE_BIN, "-i", str(src), "-f", "q+hex+lf", "--dihedrals", "--batch", "5"],
                       capture_output=True, timeout=120)
    assert r.returncode == 0, r.stderr
    _, want = run_pipe(gpu, tmp_path, frames.tobytes(), "q+hex+lf", dihedrals=True)
    assert r.stdout == want
    # stdin, and a truncated stream: exit status 5 after the whole frames
    r = subprocess.run([PIPE_BIN, "-f", "hex+lf"], input=frames[:2].tobytes() + b"\x01\x02",
                       capture_output=True, timeout=120)
    assert r.returncode == 5
    assert len(r.stdout.decode().splitlines()) == 2
    r = subprocess.run([PIPE_BIN, "-f", "hexx"], input=b"", capture_output=True, timeout=60)
    assert r.returncode == 255


def test_cached_buffers_reused(gpu, oracle_out, frames, tmp_path):
    """A run keeps its buffer set for the next run of the same shape: two runs
    in a row, a run after pipe_trim freed the cache, and runs of other shapes
    in between all write the same bytes."""
    data = frames.tobytes()
    _, first = run_pipe(gpu, tmp_path, data, "q+hex+lf", dihedrals=True, batch=4)
    _, again = run_pipe(gpu, tmp_path, data, "q+hex+lf", dihedrals=True, batch=4)
    assert again == first
    _, diag = run_pipe(gpu, tmp_path, data, "diagnostic+lf", dihedrals=True, batch=4)
    _, other = run_pipe(gpu, tmp_path, data, "q+hex+lf", dihedrals=True, batch=3)
    assert other == first
    assert gpu.pipe_trim() >= 1
    assert gpu.pipe_trim() == 0
    _, after = run_pipe(gpu, tmp_path, data, "q+hex+lf", dihedrals=True, batch=4)
    assert after == first
    _, diag2 = run_pipe(gpu, tmp_path, data, "diagnostic+lf", dihedrals=True, batch=4)
    assert diag2 == diag


def test_batch_size_throughput(gpu, tmp_path):
    """Batches of 256 frames keep at least 0.7 of the frames per second of
    batches of 64: 4096 frames from a file in the page cache, hex records to
    /dev/null.  Pinning the ring inside each call had given 0.37 (33.7k
    against 91.6k frames/s); with the ring cached, 256 measures 0.83-0.98 of
    64.  The two batch sizes alternate over two rounds (64, 256, 64, 256)
    and each keeps its best of three runs per round, so that the host's own
    drift over the test (up to 25 % between rounds) does not favour one of
    them; one run per size first sets up the cached buffers."""
    import time
    base = synthetic.images_u8(64, start=500)
    data = np.tile(base, (64, 1, 1)).tobytes()
    src = tmp_path / "frames.raw"
    src.write_bytes(data)
    rate = {64: 0.0, 256: 0.0}
    for batch in (64, 256):
        gpu.pipe(str(src), os.devnull, fmt="hex+lf", batch=batch)
    for batch in (64, 256, 64, 256):
        best = float("inf")
        for _ in range(3):
            t0 = time.perf_counter()
            n = gpu.pipe(str(src), os.devnull, fmt="hex+lf", batch=batch)
            best = min(best, time.perf_counter() - t0)
            assert n == 4096
        rate[batch] = max(rate[batch], 4096 / best)
        say(f"[pipe] batch {batch}: {4096 / best:.0f} frames/s")
    gpu.pipe_trim()
    assert rate[256] >= 0.7 * rate[64], rate
