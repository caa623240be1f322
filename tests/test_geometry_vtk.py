"""Host-side geometry reader and VTK writers of the library (no GPU needed).

  * lbm_read_geometry gives every site the numbering the reference's read_geo gives it (0 solid,
    1, 2, ... fluid in site order, added columns), pinned to read_geo itself (golden "read_geo": the
    reference's function built from its sources, SHA-256 of its int array);
  * lbm_write_vtk: ASCII files as rec_density / rec_velocity write them (header, 6 digits, 0 at solid
    sites), and the binary form the reference describes (LBM_functions.cpp:1515-1570): the same
    header with BINARY, big-endian float (or double) values, then a newline.  The values are checked
    against a sequential recomputation in Python (calcula's order).
"""
import hashlib
import os

import numpy as np
import pytest

import _pkg

Q = 19
CX = [0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1, 0, 0, 0, 0]
CY = [0, 0, 0, 1, -1, 0, 0, 1, -1, -1, 1, 0, 0, 0, 0, -1, 1, -1, 1]
CZ = [0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, -1, -1, 1, -1, 1, 1, -1]


def write_values(path, dims, values):
    nx, ny, nz = dims
    with open(path, "w") as f:
        f.write("# vtk DataFile Version 2.0\nGeometria\nASCII\nDATASET STRUCTURED_POINTS\n")
        f.write(f"DIMENSIONS {nx} {ny} {nz}\nASPECT_RATIO 1 1 1\nORIGIN 0 0 0\nPOINT_DATA {nx * ny * nz}\n")
        f.write("SCALARS Geometria float\nLOOKUP_TABLE default\n")
        for r in range(ny * nz):
            f.write(" ".join(str(v) for v in values[r * nx:(r + 1) * nx]) + " \n")


@pytest.mark.parametrize("idx", range(3))
def test_read_geometry_numbers_sites_as_read_geo(golden, tmp_path, idx):
    m = _pkg.load()
    ent = golden["read_geo"][idx]
    path = str(tmp_path / "meio.vtk")
    write_values(path, ent["dims"], ent["values"])
    nx, ny, nz, meio = m.read_geometry(path, ent["pts_in"], ent["pts_out"])
    assert [nx, ny, nz] == ent["box"]
    assert int(meio.max()) == ent["n_fluid"]
    assert hashlib.sha256(meio.astype(np.int32).tobytes()).hexdigest() == ent["meio_sha"], ent["label"]


def test_read_geometry_rejects_malformed_files(tmp_path):
    m = _pkg.load()
    with pytest.raises(m.LBMError):
        m.read_geometry(str(tmp_path / "missing.vtk"))
    p = tmp_path / "short.vtk"
    write_values(str(p), (4, 3, 1), [1] * 12)
    txt = p.read_text().rsplit(" \n", 2)[0]   # drop the last row
    p.write_text(txt)
    with pytest.raises(m.LBMError):
        m.read_geometry(str(p))


def calcula(f):
    rho = 0.0
    mx = my = mz = 0.0
    for i in range(Q):
        rho = rho + f[i]
        mx = mx + CX[i] * f[i]
        my = my + CY[i] * f[i]
        mz = mz + CZ[i] * f[i]
    if rho != 0.0:
        inv = 1.0 / rho
        return rho, (mx * inv, my * inv, mz * inv)
    return rho, (0.0, 0.0, 0.0)


def expected(fR, fB, force, meio, n):
    dens = np.zeros(n)
    vel = np.zeros((n, 3))
    for s in range(n):
        if meio is not None and meio[s] == 0:
            continue
        r = 0.0
        for i in range(Q):
            r += fR[s, i]
        dens[s] = r
        rR, vR = calcula(fR[s])
        rB, vB = calcula(fB[s])
        cR = rR / (rR + rB)
        cB = 1.0 - cR
        v = [cR * vR[a] + cB * vB[a] for a in range(3)]
        if force is not None and rR + rB > 0.0:
            v = [v[a] - 0.5 * force[s, a] / (rR + rB) for a in range(3)]
        vel[s] = v
    return dens, vel


def read_vtk(path):
    """Header lines and the values of a legacy structured-points file (ASCII or BINARY)."""
    raw = open(path, "rb").read()
    lines, pos = [], 0
    while len(lines) < 10:
        e = raw.index(b"\n", pos)
        lines.append(raw[pos:e].decode())
        pos = e + 1
        if lines[-1].startswith("VECTORS"):
            break
    kind = lines[-1].split()[-1] if lines[-1].startswith("VECTORS") else lines[-2].split()[-1]
    if lines[2] == "ASCII":
        return lines, np.array(raw[pos:].split(), dtype=np.float64), raw
    dt = ">f4" if kind == "float" else ">f8"
    body = raw[pos:]
    assert body.endswith(b"\n")
    return lines, np.frombuffer(body[:-1], dtype=dt).astype(np.float64), raw


@pytest.mark.parametrize("binary,double", [(False, False), (True, False), (True, True)])
def test_write_vtk_density_and_velocity(tmp_path, binary, double):
    m = _pkg.load()
    nx, ny = 7, 5
    n = nx * ny
    rng = np.random.default_rng(5)
    w = np.array([1 / 3] + [1 / 18] * 6 + [1 / 36] * 12)
    fR = w * (0.7 + 0.3 * rng.random((n, 1))) * (1 + 0.05 * rng.standard_normal((n, Q)))
    fB = w * (0.2 + 0.3 * rng.random((n, 1))) * (1 + 0.05 * rng.standard_normal((n, Q)))
    force = 1e-3 * rng.standard_normal((n, 3))
    meio = np.arange(1, n + 1, dtype=np.int32)
    meio[[3, 10, 11, 30]] = 0
    dens, vel = expected(fR, fB, force, meio, n)
    pd, pv = str(tmp_path / "rho_R.vtk"), str(tmp_path / "vel.vtk")
    m.write_vtk(pd, m.VTK_DENSITY, nx, ny, fR, meio=meio, binary=binary, double=double)
    m.write_vtk(pv, m.VTK_VELOCITY, nx, ny, fR, fB, force, meio=meio, binary=binary, double=double)
    typ = "double" if (double or not binary) else "float"
    form = "BINARY" if binary else "ASCII"
    hd, vd, _ = read_vtk(pd)
    assert hd == ["# vtk DataFile Version 2.0", "Densidade", form, "DATASET STRUCTURED_POINTS", f"DIMENSIONS {nx} {ny} 1",
                  "ASPECT_RATIO 1 1 1", "ORIGIN 0 0 0", f"POINT_DATA {n}", f"SCALARS densidade {typ}",
                  "LOOKUP_TABLE default"]
    hv, vv, _ = read_vtk(pv)
    assert hv == ["# vtk DataFile Version 2.0", "Velocidade", form, "DATASET STRUCTURED_POINTS", f"DIMENSIONS {nx} {ny} 1",
                  "ASPECT_RATIO 1 1 1", "ORIGIN 0 0 0", f"POINT_DATA {n}", f"VECTORS velocidade {typ}"]
    vv = vv.reshape(n, 3)
    if not binary:   # the stream's default 6 significant digits
        np.testing.assert_allclose(vd, dens, rtol=6e-6, atol=0)
        np.testing.assert_allclose(vv, vel, rtol=6e-6, atol=1e-12)
    elif double:
        assert np.array_equal(vd, dens) and np.array_equal(vv, vel)
    else:
        assert np.array_equal(vd, dens.astype(np.float32)) and np.array_equal(vv, vel.astype(np.float32))
    assert np.all(vd[meio == 0] == 0) and np.all(vv[meio == 0] == 0)


def test_write_vtk_rejects_bad_arguments(tmp_path):
    m = _pkg.load()
    f = np.zeros(4 * Q)
    with pytest.raises(m.LBMError):
        m.write_vtk(str(tmp_path / "x.vtk"), 7, 2, 2, f)
    with pytest.raises(m.LBMError):
        m.write_vtk(str(tmp_path / "x.vtk"), m.VTK_VELOCITY, 2, 2, f)   # no blue populations
    with pytest.raises(m.LBMError):
        m.write_vtk(os.path.join(str(tmp_path), "no", "dir", "x.vtk"), m.VTK_DENSITY, 2, 2, f)
