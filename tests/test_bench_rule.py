"""bench.py's choice of `value`: a learned mode counts only when it is at
least as accurate as the plain run (steps over the tolerance, worst step
error, energy peak with 5 % slack), and the fastest mode that counts wins."""
import importlib.util
import os

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
_spec = importlib.util.spec_from_file_location("bench_mod", os.path.join(ROOT, "bench.py"))
bench = importlib.util.module_from_spec(_spec)
_spec.loader.exec_module(bench)


def _mode(value, nonconv, worst, emax=500.0):
    return {"value": value, "nonconverged_steps": nonconv, "worst_density_error": worst,
            "energy_J": {"start": 500.0, "max": emax, "end": emax}}


def test_plain_without_weights():
    assert bench.choose_value(_mode(1.0, 0, 0.01), {"learned": None})[0] == "plain"


def test_fastest_accurate_mode_wins():
    plain = _mode(100.0, 10, 0.05)
    modes = {"learned": _mode(110.0, 10, 0.05), "learned_always": _mode(120.0, 8, 0.04)}
    assert bench.choose_value(plain, modes)[0] == "learned_always"
    modes = {"learned": _mode(130.0, 10, 0.05), "learned_always": _mode(120.0, 8, 0.04)}
    assert bench.choose_value(plain, modes)[0] == "learned"


def test_each_accuracy_rule_excludes_a_mode():
    plain = _mode(100.0, 10, 0.05)
    for bad in (_mode(200.0, 11, 0.05), _mode(200.0, 10, 0.051), _mode(200.0, 10, 0.05, emax=526.0)):
        mode, why = bench.choose_value(plain, {"learned": bad})
        assert mode == "plain" and "learned:" in why
    # within the energy slack, faster: counts
    assert bench.choose_value(plain, {"learned": _mode(200.0, 10, 0.05, emax=524.0)})[0] == "learned"


def test_accurate_but_slower_is_not_value():
    mode, why = bench.choose_value(_mode(100.0, 10, 0.05), {"learned": _mode(90.0, 5, 0.01)})
    assert mode == "plain" and "slower" in why
