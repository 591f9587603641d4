"""The .xyz reader (owm_xyz_parse): every parsed double equals Python's
float() of the same token (correctly rounded, like strtod and fscanf), in
every text layout the reference's readXYZfile accepts; the header, blank
lines, the host-parsed lines and the errors; and a parsed cloud runs through
the pipeline to the oracle's seeds."""
import os

import numpy as np
import pytest

import pyoracle as O
from fastowm_amd import owm, synth

pytestmark = pytest.mark.gpu


def say(*a):
    print("[case]", *a, flush=True)


def expect(text):
    """Reference values: the first three tokens of every non-blank line."""
    xs, ys, zs = [], [], []
    for line in text.split("\n"):
        tok = line.split()
        if not tok:
            continue
        xs.append(float(tok[0]))
        ys.append(float(tok[1]))
        zs.append(float(tok[2]))
    return (np.array(xs, np.float64), np.array(ys, np.float64), np.array(zs, np.float64))


def same_bits(a, b):
    return a.shape == b.shape and np.array_equal(a.view(np.uint64), b.view(np.uint64))


def check(ctx, text, host_lines=None):
    c = ctx.parse_xyz(text.encode())
    got = c.to_numpy()
    want = expect(text)
    assert c.n == len(want[0])
    for g, w in zip(got, want):
        assert same_bits(g, w)
    if host_lines is not None:
        assert c.host_lines == host_lines
    return c


def cloud_text(n, seed, fmt, sep=" ", eol="\n", extra=""):
    t = synth.scaled(synth.config_c1(seed), n)
    x, y, z = t.generate()
    rows = [fmt % (a, b, c) for a, b, c in zip(x, y, z)]
    return eol.join(r.replace(" ", sep) + extra for r in rows) + eol


@pytest.fixture(scope="module")
def ctx():
    c = owm.Context(0)
    yield c
    c.close()


@pytest.mark.parametrize("fmt,sep,eol,extra", [
    ("%.2f %.2f %.2f", " ", "\n", ""),
    ("%.3f %.3f %.3f", "\t", "\r\n", ""),
    ("%+.2f %.4f %.1f", " ", "\n", " 123 45 67 89"),
    ("%.6e %.6e %.6e", "  ", "\n", ""),
])
def test_fast_path_layouts(ctx, fmt, sep, eol, extra):
    say("layout", fmt, repr(sep), repr(eol), repr(extra))
    text = cloud_text(30_000, 7, fmt, sep, eol, extra)
    check(ctx, text, host_lines=0)
    check(ctx, text.rstrip("\r\n"), host_lines=0)  # no final newline


def test_host_parsed_lines(ctx):
    say("round-trip digits and odd tokens")
    text = cloud_text(20_000, 8, "%.17g %.17g %.17g")
    c = check(ctx, text)
    assert c.host_lines > 0
    odd = ("1.5e2 -0.00 .5\n5. +7 1e-5\n123456789012345678901.5 2 3\n"
           "inf -inf nan\n0.000000000000000000000000001 1e300 -2.5E-3\n"
           "4503599627370497 9007199254740993 1e23\n")
    check(ctx, odd)


def test_blank_lines_and_header(ctx):
    say("blank lines")
    body = cloud_text(5_000, 9, "%.2f %.2f %.2f")
    lines = body.split("\n")
    lines[10] = ""
    lines[20] = "   \t "
    lines.insert(0, "")
    text = "\n".join(lines) + "\n\n"
    check(ctx, text)
    say("header")
    x, y, z = expect(body)
    box = (x.min(), x.max(), y.min(), y.max())
    hdr = "%d\n%.2f\n%.2f\n%.2f\n%.2f\n" % ((len(x) - 3,) + box)
    c = ctx.parse_xyz((hdr + body).encode())
    assert c.bbox == tuple(float("%.2f" % v) for v in box)
    assert c.n == len(x) - 3  # the reference reads N points
    gx, gy, gz = c.to_numpy()
    assert same_bits(gx, x[:-3]) and same_bits(gz, z[:-3])
    with pytest.raises(owm.OWMError) as e:
        ctx.parse_xyz(("%d\n1\n2\n3\n4\n" % (len(x) + 1) + body).encode())
    assert e.value.name == "E_ARG"


def test_every_line_on_the_host(ctx):
    say("list capacity exceeded")
    text = cloud_text(3_000, 10, "%.17g %.17g %.17g")
    os.environ["OWM_XYZ_LIST_CAP"] = "100"
    try:
        c = check(ctx, text)
        assert c.host_lines == 3_000
        lines = text.split("\n")
        for k in range(0, 400, 2):
            lines[k] = " "
        check(ctx, "\n".join(lines))
    finally:
        del os.environ["OWM_XYZ_LIST_CAP"]


def test_retired_cloud_raises(ctx):
    say("retired cloud")
    a = ctx.parse_xyz(b"1 2 3\n4 5 6\n")
    b = ctx.parse_xyz(b"7 8 9\n")
    with pytest.raises(owm.OWMError):
        a.to_numpy()
    with pytest.raises(owm.OWMError):
        ctx.run_cloud(a)
    assert same_bits(b.to_numpy()[2], np.array([9.0]))


def test_malformed_line(ctx):
    say("two numbers on a line")
    with pytest.raises(owm.OWMError) as e:
        ctx.parse_xyz(b"1 2 3\n4 5\n6 7 8\n")
    assert e.value.name == "E_ARG"
    assert "line 2" in ctx.last_error()
    assert ctx.parse_xyz(b"").n == 0


def test_parsed_cloud_runs_to_the_oracle(ctx):
    say("headed c1 text -> seeds")
    t = synth.scaled(synth.config_c1(seed=3), 400_000)
    x, y, z = t.generate()
    rows = "\n".join("%.2f %.2f %.2f" % r for r in zip(x, y, z)) + "\n"
    px, py, pz = expect(rows)
    box = (float(px.min()), float(px.max()), float(py.min()), float(py.max()))
    hdr = "%d\n%.17g\n%.17g\n%.17g\n%.17g\n" % ((len(px),) + box)
    c = ctx.parse_xyz((hdr + rows).encode())
    assert c.host_lines == 0 and c.bbox == box
    got = ctx.run_cloud(c, 12, 20, 0.8, 32)
    want, _, _ = O.run(px, py, pz, 12, 20, 0.8, 32, box=box)
    assert np.array_equal(got, want)
    assert np.array_equal(ctx.run_tile(px, py, pz, 12, 20, 0.8, 32, bbox=box), want)
