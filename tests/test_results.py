"""Result writers (host code of libmantaray_hip.so; no GPU).

RayResult JSON (ray_result.rs:74-151): the reference's known answers
(ray_result.rs:171-205: one row, and the NaN-row truncation), the empty
result, values reading back bitwise, and serde_json's f64 layout (the ryu
crate's: ".0" on integral values, decimal for exponents in (-5, 16],
scientific outside).  TSV (ray.rs:218-243): header, rows up to the first NaN
x, trailing spaces, END, append mode, Rust Display's f64 text (no exponent,
no ".0"), values reading back bitwise.
"""
import json

import numpy as np
import pytest

from mantaray_amd import results


def rows(*r):
    return np.array(r, dtype=np.float64)


def test_reference_known_answers():
    s = results.ray_result_json(rows([1.0, 2.0, 3.0, 4.0, 5.0]))
    assert s == '{"t_vec":[1.0],"x_vec":[2.0],"y_vec":[3.0],"kx_vec":[4.0],"ky_vec":[5.0]}'
    nan = np.nan
    s = results.ray_result_json(rows([0.0, 1.0, 1.0, 1.0, 1.0], [1.0, 1.0, nan, nan, 1.0],
                                     [2.0, nan, nan, nan, nan], [3.0, 2.0, 2.0, 2.0, 2.0]))
    assert s == '{"t_vec":[0.0],"x_vec":[1.0],"y_vec":[1.0],"kx_vec":[1.0],"ky_vec":[1.0]}'
    s = results.ray_result_json(np.zeros((0, 5)))
    assert s == '{"t_vec":[],"x_vec":[],"y_vec":[],"kx_vec":[],"ky_vec":[]}'


@pytest.mark.parametrize("v,text", [
    (123.0, "123.0"), (0.1, "0.1"), (-2.5, "-2.5"), (1e15, "1000000000000000.0"),
    (1e16, "1e16"), (1.5e16, "1.5e16"), (12345678901234567.0, "1.2345678901234568e16"),
    (1e-5, "0.00001"), (1.25e-5, "0.0000125"), (1e-6, "1e-6"), (3.5e-7, "3.5e-7"),
    (5e-324, "5e-324"), (1.7976931348623157e308, "1.7976931348623157e308"),
    (-0.0, "-0.0"), (0.0, "0.0"), (9999999999999998.0, "9999999999999998.0"),
])
def test_json_number_layout(v, text):
    s = results.ray_result_json(rows([v, 0.0, 0.0, 0.0, 0.0]))
    assert s.startswith('{"t_vec":[' + text + "]"), s


def test_json_round_trip():
    rng = np.random.default_rng(0)
    n = 3000
    vals = rng.standard_normal((n, 5)) * 10.0 ** rng.integers(-320, 300, (n, 5))
    vals[:, 1:][np.abs(vals[:, 1:]) < 1e-300] = 1.0
    s = results.ray_result_json(vals)
    d = json.loads(s)
    for i, k in enumerate(("t_vec", "x_vec", "y_vec", "kx_vec", "ky_vec")):
        assert np.array_equal(np.array(d[k]), vals[:, i]), k


def test_json_from_trace_result():
    t = np.arange(4.0)
    x = np.array([[0.0, 10.0], [1.0, 11.0], [np.nan, 12.0], [np.nan, 13.0]])
    r = {"t": t, "x": x, "y": x + 0.5, "kx": x * 2, "ky": -x}
    d = json.loads(results.ray_result_json(r, 0))
    assert d["x_vec"] == [0.0, 1.0] and d["t_vec"] == [0.0, 1.0]
    d = json.loads(results.ray_result_json(r, 1))
    assert d["x_vec"] == [10.0, 11.0, 12.0, 13.0] and d["ky_vec"] == [-10.0, -11.0, -12.0, -13.0]


def test_tsv_layout_and_append(tmp_path):
    p = tmp_path / "out.tsv"
    r = rows([0.0, 1.0, -0.0, 1e-7, 1e21], [0.5, 0.1, 2.0, -3.25, np.nan],
             [1.0, np.nan, 0.0, 0.0, 0.0], [2.0, 9.0, 9.0, 9.0, 9.0])
    results.append_tsv(p, r)
    results.append_tsv(p, rows([7.0, 8.0, 9.0, 10.0, 11.0]))
    text = p.read_text()
    assert text == ("t x y kx ky\n"
                    "0 1 -0 0.0000001 1000000000000000000000  \n"
                    "0.5 0.1 2 -3.25 NaN  \n"
                    "END\n"
                    "t x y kx ky\n"
                    "7 8 9 10 11  \n"
                    "END\n")


def test_tsv_round_trip(tmp_path):
    rng = np.random.default_rng(1)
    vals = rng.standard_normal((200, 5)) * 10.0 ** rng.integers(-30, 30, (200, 5))
    p = tmp_path / "rt.tsv"
    results.append_tsv(p, vals)
    lines = p.read_text().split("\n")
    assert lines[0] == "t x y kx ky" and lines[-2] == "END"
    got = np.array([[float(tok) for tok in ln.split()] for ln in lines[1:-2]])
    assert np.array_equal(got, vals)
    assert all(ln.endswith("  ") and "e" not in ln for ln in lines[1:-2])


def test_writer_errors(tmp_path):
    from mantaray_amd import _capi as C
    with pytest.raises(C.MantarayError):
        results.append_tsv(tmp_path / "missing_dir" / "x.tsv", rows([0, 1, 2, 3, 4]))
