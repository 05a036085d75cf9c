"""Configuration text files (IO/SaveConfiguration.cpp:3-63, IO/LoadConfiguration.cpp:3-144) through
the C ABI: the line format of the reference, bit-exact FP64 round trip, files written elsewhere in
that format (any line order, fewer digits), and the count check of the loader."""
import numpy as np
import pytest

from overlap_amd import configs as cf
from overlap_amd import ovl

pytestmark = pytest.mark.gpu


def gauge(nc, n, seed):
    g = cf.rng(seed)
    links = cf.su2_links(n, 0.5, g) if nc == 2 else cf.su3_links(n, 0.5, g)
    return links, cf.efield(n, nc, 0.3, g)


@pytest.mark.parametrize("nc", [2, 3])
def test_save_load_round_trip_and_format(gpu, nc, tmp_path):
    n = (5, 4, 3)
    V = 60
    lr, ng = (4, 3) if nc == 2 else (18, 8)
    links, e = gauge(nc, n, 40 + nc)
    uf, ef = tmp_path / "u.txt", tmp_path / "e.txt"
    with ovl.Context(nc, n, 4) as ctx:
        ctx.set_links(links)
        ctx.set_efield(e)
        ctx.save_configuration(uf, ef)
    ul = uf.read_text().splitlines()
    el = ef.read_text().splitlines()
    assert len(ul) == len(el) == 3 * V
    k = 0
    for z in range(n[2]):
        for y in range(n[1]):
            for x in range(n[0]):
                for mu in range(3):
                    tu, te = ul[k].split(), el[k].split()
                    assert [int(v) for v in tu[:4]] == [x, y, z, mu] == [int(v) for v in te[:4]]
                    l = mu + 3 * (x + n[0] * (y + n[1] * z))
                    assert len(tu) == 4 + lr and len(te) == 4 + ng
                    assert ul[k].endswith(" ")  # MatrixToString puts a space after every value
                    assert np.array_equal(np.array(tu[4:], float), links[l * lr:(l + 1) * lr])
                    assert np.array_equal(np.array(te[4:], float), e[l * ng:(l + 1) * ng])
                    k += 1
    with ovl.Context(nc, n, 4) as ctx:
        ctx.load_configuration(uf, ef)
        assert np.array_equal(ctx.get_links(), links)
        assert np.array_equal(ctx.get_efield(), e)


def test_load_foreign_file_any_order_and_errors(gpu, tmp_path):
    n = (4, 3, 3)
    V = 36
    links, e = gauge(2, n, 77)
    rows = []
    for s in range(V):
        x, y, z = s % 4, (s // 4) % 3, s // 12
        for mu in range(3):
            l = mu + 3 * s
            rows.append(f"{x} {y} {z} {mu} " + " ".join(f"{v:.6f}" for v in links[4 * l:4 * l + 4]))
    order = cf.rng(5).permutation(len(rows))
    uf = tmp_path / "u.txt"
    uf.write_text("\n".join(rows[i] for i in order) + "\n\n")
    want = np.array([float(f"{v:.6f}") for v in links])
    with ovl.Context(2, n, 4) as ctx:
        ctx.load_configuration(uf)
        assert np.array_equal(ctx.get_links(), want)
        bad = tmp_path / "bad.txt"
        bad.write_text("\n".join(rows[:-1]) + "\n")  # one link missing
        with pytest.raises(ovl.OvlError) as ex:
            ctx.load_configuration(bad)
        assert ex.value.code == ovl.E_ARG
        bad.write_text("\n".join(rows + rows[:1]) + "\n")  # one link twice
        with pytest.raises(ovl.OvlError) as ex:
            ctx.load_configuration(bad)
        assert ex.value.code == ovl.E_ARG
        bad.write_text(rows[0].replace("0 0 0 0", "4 0 0 0", 1) + "\n")  # site outside the lattice
        with pytest.raises(ovl.OvlError) as ex:
            ctx.load_configuration(bad)
        assert ex.value.code == ovl.E_ARG
        # the failed loads left the links unchanged
        assert np.array_equal(ctx.get_links(), want)


def test_ensemble_system_files(gpu, tmp_path):
    dims = [(4, 4, 3), (3, 5, 4)]
    ens = ovl.Ensemble(2, dims, 4)
    data = [gauge(2, d, 90 + k) for k, d in enumerate(dims)]
    for k, (links, e) in enumerate(data):
        ens.set_system_links(k, links)
        ens.set_system_efield(k, e)
    ens.save_configuration(tmp_path / "u1.txt", tmp_path / "e1.txt", s=1)
    ens2 = ovl.Ensemble(2, dims[::-1], 4)
    ens2.load_configuration(tmp_path / "u1.txt", tmp_path / "e1.txt", s=0)
    assert np.array_equal(ens2.get_system_links(0), data[1][0])
    assert np.array_equal(ens2.get_system_efield(0), data[1][1])
    with pytest.raises(ovl.OvlError):  # a 3x5x4 file does not fit a 4x4x3 system
        ens2.load_configuration(tmp_path / "u1.txt", None, s=1)
