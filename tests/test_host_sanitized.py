"""Host code under AddressSanitizer and UndefinedBehaviorSanitizer: the library's host
sources (SBVH builder, cache, generators, RGBE / OBJ / JSON loaders; no device code)
linked with tests/hostcheck/host_fuzz.cpp and driven with valid files and with
truncated, bit-flipped and malformed copies. Every load must end with a status
(an error for a bad file), never with a sanitizer report or a crash."""
import concurrent.futures as cf
import importlib.util
import json
import os
import shutil
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
FLAGS = ["-O1", "-g", "-std=c++17", "-fsanitize=address,undefined", "-fno-sanitize-recover=all",
         "-fno-omit-frame-pointer", "-I", os.path.join(ROOT, "include")]
ENV = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:allocator_may_return_null=1:abort_on_error=0",
           UBSAN_OPTIONS="print_stacktrace=1")


def _host_sources():
    spec = importlib.util.spec_from_file_location("_tt_build", os.path.join(ROOT, "trenchanttracer_amd", "build.py"))
    b = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(b)
    return [os.path.join(b.CSRC, s) for s in b.HOST_SRC]


@pytest.fixture(scope="module")
def fuzz(tmp_path_factory):
    if shutil.which("g++") is None:
        pytest.skip("g++ not available")
    d = tmp_path_factory.mktemp("hostcheck")
    srcs = _host_sources() + [os.path.join(ROOT, "tests", "hostcheck", "host_fuzz.cpp")]

    def cc(src):
        obj = str(d / (os.path.basename(src) + ".o"))
        r = subprocess.run(["g++"] + FLAGS + ["-c", src, "-o", obj], capture_output=True, text=True)
        assert r.returncode == 0, r.stderr[-3000:]
        return obj

    with cf.ThreadPoolExecutor(max_workers=min(8, os.cpu_count() or 2)) as ex:
        objs = list(ex.map(cc, srcs))
    exe = str(d / "host_fuzz")
    r = subprocess.run(["g++"] + FLAGS + objs + ["-o", exe, "-lpthread"], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-3000:]

    def run(*args):
        r = subprocess.run([exe] + [str(a) for a in args], capture_output=True, text=True, env=ENV, timeout=600)
        bad = r.returncode != 0 or "runtime error" in r.stderr or "Sanitizer" in r.stderr
        assert not bad, f"exit {r.returncode}\n{r.stderr[-4000:]}"
        return r.stdout

    return run


def _mutants(data, rng, n_flips=24):
    """Truncations at assorted lengths and copies with 1-8 bytes overwritten."""
    out = []
    for k in sorted({0, 1, 2, 5, 10, 27, 28, 29, 64, len(data) // 4, len(data) // 2, len(data) - 1}):
        if 0 <= k < len(data):
            out.append(data[:k])
    for _ in range(n_flips):
        b = bytearray(data)
        for _ in range(int(rng.integers(1, 9))):
            b[int(rng.integers(0, len(b)))] = int(rng.integers(0, 256))
        out.append(bytes(b))
    return out


def _write_all(d, stem, blobs):
    paths = []
    for i, b in enumerate(blobs):
        p = d / f"{stem}_{i}"
        p.write_bytes(b)
        paths.append(p)
    return paths


def test_generated_files_round_trip_clean(fuzz, tmp_path):
    out = fuzz("gen", tmp_path)
    assert "gen ok" in out
    out = fuzz("load", "bvh", tmp_path / "bvh.cache")
    assert out.strip().endswith(" 0")            # TT_OK


def test_corrupt_bvh_caches(fuzz, tmp_path):
    fuzz("gen", tmp_path)
    data = (tmp_path / "bvh.cache").read_bytes()
    rng = np.random.default_rng(1)
    blobs = _mutants(data, rng)
    # header words (7 x u32) set one at a time to huge or zero counts
    for w in range(7):
        for v in (0, 0x7FFFFFFF, 0xFFFFFFF0):
            b = bytearray(data)
            b[4 * w:4 * w + 4] = int(v).to_bytes(4, "little")
            blobs.append(bytes(b))
    paths = _write_all(tmp_path, "cache", blobs)
    out = fuzz("load", "bvh", *paths)
    statuses = [int(l.split()[-1]) for l in out.strip().splitlines()]
    assert len(statuses) == len(paths)
    assert all(s != 0 for s in statuses[:12])     # every truncation is refused


def test_corrupt_hdr_files(fuzz, tmp_path):
    fuzz("gen", tmp_path)
    data = (tmp_path / "sky.hdr").read_bytes()
    rng = np.random.default_rng(2)
    blobs = _mutants(data, rng)
    head = data[:data.index(b"-Y")]
    for res in (b"-Y 0 +X 5\n", b"-Y 5 +X -3\n", b"-Y 99999999 +X 99999999\n", b"+Y 9 +X 17\n", b"-Y 9\n"):
        blobs.append(head + res + data[data.index(b"\n", data.index(b"-Y")) + 1:])
    blobs.append(b"#?RADIANCE\nFORMAT=32-bit_rle_rgbe\n\n-Y 9 +X 17\n" + bytes(range(256)) * 4)
    paths = _write_all(tmp_path, "hdr", blobs)
    out = fuzz("load", "hdr", *paths)
    assert len(out.strip().splitlines()) == len(paths)


OBJ = ("mtllib lib.json\n"
       "v -1 -1 -1\nv 1 -1 -1\nv 1 1 -1\nv -1 1 -1\nv -1 -1 1\nv 1 -1 1\nv 1 1 1\nv -1 1 1\n"
       "vn 0 0 1\nvn 0 1 0\n"
       "g sides\nusemtl body\nf 1//1 2//1 6//2 5//2\nf 2 3 7 6\nf 3/1/1 4/1/1 8/1/2 7/1/2\n"
       "g caps\nusemtl trim\nf 1 4 3 2\nf 5 6 7 8\n")
MTL = {"body": {"type": "Diff", "diffuseColor": [0.8, 0.7, 0.6]},
       "trim": {"type": "Coat", "diffuseColor": [0.3, 0.3, 0.6], "specularColor": [0.9, 0.9, 0.9]}}


def test_malformed_obj_files(fuzz, tmp_path):
    (tmp_path / "lib.json").write_text(json.dumps(MTL))
    (tmp_path / "ok.obj").write_text(OBJ)
    out = fuzz("load", "obj", tmp_path / "ok.obj")
    assert out.strip().endswith(" 0")
    rng = np.random.default_rng(3)
    blobs = _mutants(OBJ.encode(), rng)
    bad_lines = ["f 1 2 99999", "f -1 -2 -3", "f -99 1 2", "f 0 0 0", "f 1", "f 1 2", "f 1/2/3/4 2 3",
                 "f a b c", "v nan inf -inf", "v 1e39 -1e39 0", "v 1", "vn", "usemtl", "usemtl nosuch",
                 "mtllib missing.json", "g", "f 1//99 2//99 3//99", "f 1/1/1 2/1/1",
                 "f " + " ".join(str(i % 8 + 1) for i in range(5000)), "x" * 1_000_000]
    for line in bad_lines:
        blobs.append((OBJ + line + "\n").encode())
        blobs.append((line + "\n" + OBJ).encode())
    blobs.append(bytes(rng.integers(0, 256, 4096, dtype=np.uint8)))
    paths = _write_all(tmp_path, "obj", blobs)
    out = fuzz("load", "obj", *paths)
    assert len(out.strip().splitlines()) == len(paths)


def test_malformed_configs(fuzz, tmp_path, scene_files):
    base = scene_files.parent
    out = fuzz("load", "config", scene_files)
    assert out.strip().endswith(" 0")
    text = scene_files.read_text()
    cfg = json.loads(text)
    rng = np.random.default_rng(4)
    blobs = _mutants(text.encode(), rng)
    variants = [
        {**cfg, "objects": "nope"},
        {**cfg, "objects": [{"type": "Sphere"}]},
        {**cfg, "objects": [{"type": "Plane", "p0": [1, 2], "material": "diff1"}]},
        {**cfg, "objects": [{"type": "Mesh", "file": "missing.obj"}]},
        {**cfg, "objects": [{"type": "Mesh", "file": "scene.json"}]},
        {**cfg, "objects": [{"type": "Sphere", "position": [0, 0, 0], "radius": 1, "material": "nosuch"}]},
        {**cfg, "objects": [{"type": "Mesh", "file": "thing.obj", "transform": {"scale": "big"}}]},
        {**cfg, "material-file": "missing.json"},
        {**cfg, "material-file": "scene.json"},
        {**cfg, "camera": {"width": -5, "height": 1e12}},
        {**cfg, "camera": None},
        [], 7, "text", None,
    ]
    for v in variants:
        blobs.append(json.dumps(v).encode())
    blobs.append(b"[" * 100000 + b"]" * 100000)
    blobs.append(b"{\"a\":" * 50000 + b"1" + b"}" * 50000)
    blobs.append(b"\"" + b"\\u12" + b"\"")
    blobs.append(b"{\"objects\": [1e999999, -1e999999, 123456789012345678901234567890]}")
    paths = []
    for i, b in enumerate(blobs):
        p = base / f"cfg_{i}.json"
        p.write_bytes(b)
        paths.append(p)
    out = fuzz("load", "config", *paths)
    assert len(out.strip().splitlines()) == len(paths)
