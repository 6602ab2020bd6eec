"""Host-only output formats of the C ABI: the .aln file (CSaveAlign /
CLoadAlign text format) and the Relion .star file (CGenStarFile), checked
line by line against the reference's printf formats restated here."""
import numpy as np

from motioncor3_amd import lib as L


def _aln_lines(sx, sy, nstk, px, py, t0, t1, g, ps, pc):
    # Align/CSaveAlign.cpp:84-151
    out = ["setting", "   stackSize: %d %d %d %d" % (sx, sy, nstk, 1),
           "   patches: %d %d %d" % (px, py, 0), "   throw: %d %d" % (t0, t1),
           "#    Global Shift", "#---------------------", "globalShift", "   stackID 0"]
    out += ["%6d  %+8.3f  %+8.3f" % (i, g[i, 0], g[i, 1]) for i in range(len(g))]
    if ps is not None:
        P = pc.shape[0]
        out += ["#    Local Shift", "#---------------------", "localShift"]
        for q in range(P):
            out += ["   patchID: %d" % q, "   Converge: %d" % 1]
            out += ["%6d  %8.2f  %8.2f  %+8.3f  %+8.3f" % (i, pc[q, 0], pc[q, 1], ps[i, q, 0], ps[i, q, 1])
                    for i in range(len(g))]
            if q < P - 1:
                out.append("")
    return out


def test_aln_global_round_trip(tmp_path):
    rng = np.random.default_rng(1)
    N = 12
    g = rng.normal(0, 5, (N, 2)).astype(np.float32)
    p = L.default_params(throw0=1)
    path = str(tmp_path / "a.aln")
    L.save_aln(path, p, 4096, 4096, N + 3, g)
    lines = open(path).read().split("\n")
    assert lines[:-1] == _aln_lines(4096, 4096, N + 3, 0, 0, 1, 2, g.astype(np.float64), None, None)
    back = L.load_aln(path)
    assert (back["sx"], back["sy"], back["n"], back["patch"]) == (4096, 4096, N, (0, 0))
    want = np.array([[float("%+8.3f" % v) for v in row] for row in g], np.float32)
    assert np.array_equal(back["global_shift"], want)


def test_aln_patch_round_trip(tmp_path):
    rng = np.random.default_rng(2)
    N, px, py = 7, 3, 2
    P = px * py
    g = rng.normal(0, 5, (N, 2)).astype(np.float32)
    ps = rng.normal(0, 1, (N, P, 2)).astype(np.float32)
    pc = rng.uniform(0, 4000, (P, 2)).astype(np.float32)
    p = L.default_params(patch=(px, py))
    path = str(tmp_path / "b.aln")
    L.save_aln(path, p, 5760, 4092, N, g, ps, pc)
    lines = open(path).read().split("\n")
    assert lines[:-1] == _aln_lines(5760, 4092, N, px, py, 0, 0, g.astype(np.float64),
                                     ps.astype(np.float64), pc.astype(np.float64))
    back = L.load_aln(path)
    f3 = np.vectorize(lambda v: float("%+8.3f" % v))
    f2 = np.vectorize(lambda v: float("%8.2f" % v))
    assert np.array_equal(back["patch_shift"], f3(ps).astype(np.float32))
    assert np.array_equal(back["patch_center"], f2(pc).astype(np.float32))
    assert np.array_equal(back["global_shift"], f3(g).astype(np.float32))


def test_star_file(tmp_path):
    rng = np.random.default_rng(3)
    N = 5
    g = np.round(rng.normal(0, 5, (N, 2)), 2).astype(np.float32)
    p = L.default_params(pix_size=0.83, fm_dose=1.2, kv=300, movie_name="/data/m/movie_0001.mrc",
                         throw0=2)
    path = str(tmp_path / "m.star")
    L.save_star(path, p, 5760, 4092, 52, g, gain_name="/refs/gain.mrc")
    # CGenStarFile.cpp:39-120
    want = ["", "# version 30001", "", "data_general", "",
            "%40s %-16d" % ("_rlnImageSizeX", 5760), "%40s %-16d" % ("_rlnImageSizeY", 4092),
            "%40s %-16d" % ("_rlnImageSizeZ", 52), "%40s %s" % ("_rlnMicrographMovieName", "movie_0001.mrc"),
            "%40s %s" % ("_rlnMicrographGainName", "gain.mrc"),
            "%40s %-16.6f" % ("_rlnMicrographOriginalPixelSize", np.float32(0.83)),
            "%40s %-16.6f" % ("rlnMicrographBinning", 1.0),
            "%40s %-16.6f" % ("_rlnMicrographDoseRate", np.float32(1.2)),
            "%40s %-16.6f" % ("_rlnMicrographPreExposure", 0.0),
            "%40s %-16.6f" % ("_rlnVoltage", 300.0),
            "%40s %-16d" % ("_rlnMicrographStartFrame", 3),
            "%40s %-16d" % ("_rlnMotionModelVersion", 0), "",
            "# version 30001", "", "data_global_shift", "", "loop",
            "_rlnMicrographFrameNumber #1", "_rlnMicrographShiftX #2", "_rlnMicrographShiftY #3"]
    want += ["%13d %13.6f %13.6f" % (i + 1, np.float32(g[i, 0] - g[0, 0]), np.float32(g[i, 1] - g[0, 1]))
             for i in range(N)]
    want += [""]
    assert open(path).read().split("\n")[:-1] == want
