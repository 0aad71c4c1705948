"""The OVITO .data writer (src/io/ff_ovito.jl:81-116) on the CPU: file layout,
Julia's float spelling, exact round trips.

The digits are the shortest round-trip ones (pinned below by round trips);
Julia's exact choice between plain and exponent notation is restated from its
printing rules and checked on the documented cases only, since no Julia and no
.data fixture of the reference exist here: the spelling is parity unpinned."""
import numpy as np

from peridyn_amd import snapshot as S


def test_julia_float_spelling():
    cases = [(1.0, "1.0"), (0.1, "0.1"), (-2.5, "-2.5"), (1e-5, "1.0e-5"), (0.0001, "0.0001"),
             (1e16, "1.0e16"), (1e15, "1000000000000000.0"), (1.5e-300, "1.5e-300"), (0.0, "0.0"),
             (-0.0, "-0.0"), (float("nan"), "NaN"), (float("inf"), "Inf"), (float("-inf"), "-Inf"),
             (5e-324, "5.0e-324")]
    for x, want in cases:
        assert S.julia_float(x) == want, (x, S.julia_float(x))


def test_julia_float_round_trips():
    rng = np.random.default_rng(7)
    xs = np.concatenate([rng.standard_normal(2000) * 10.0 ** rng.integers(-20, 20, 2000),
                         rng.integers(-10**6, 10**6, 200).astype(float)])
    for x in xs:
        assert float(S.julia_float(float(x))) == float(x)


def test_write_ovito_layout_and_round_trip(tmp_path):
    N = 5
    rng = np.random.default_rng(3)
    pos = rng.standard_normal((3, N))
    cols = dict(id=np.arange(1, N + 1), type=np.array([1, 1, 2, 2, 99]), position=pos,
                velocity=rng.standard_normal((3, N)), acceleration=rng.standard_normal((3, N)) * 1e6,
                mass=rng.uniform(size=N), volume=np.ones(N), damage=np.array([0.0, 0.25, 0.5, 1.0, 0.0]))
    fn = tmp_path / "s.data"
    S.write_ovito(str(fn), **cols)
    lines = open(fn).read().split("\n")
    assert lines[0] == f"{N} "
    assert lines[1] == "# acceleration | damage | id | mass | position | type | velocity | volume"
    first = lines[2].split(" ")
    assert lines[2].endswith(" ") and len(first) == 3 + 1 + 1 + 1 + 3 + 1 + 3 + 1 + 1
    assert first[4] == "1" and first[9] == "1"  # id and type print as integers
    back = S.read_ovito(str(fn))
    for k, v in cols.items():
        assert np.array_equal(back[k], v), k
