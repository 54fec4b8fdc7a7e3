"""bench.py's JSON contract on the CPU: --gpus above 1 is refused with a
message, and the summary gathers every leg's speedup and worst error."""
import importlib.util
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def load_bench():
    spec = importlib.util.spec_from_file_location("bench", os.path.join(ROOT, "bench.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def test_summary_gathers_the_legs():
    b = load_bench()
    line = {"speedup_vs_fp64_gpu": 3.2, "speedup_vs_cpu": 4800.0,
            "rel_err_vs_fp64": {"forward": 6e-8, "adjoint": 3e-7},
            "prism": {"speedup_vs_fp64_gpu": 220.0, "speedup_vs_fp64_same_algorithm": 3.1,
                      "rel_err_vs_fp64": {"forward": 4e-7, "adjoint": 5e-7}},
            "batched": {"split3": {"rel_err_vs_fp64": 1.1e-6,
                                   "adjoint": {"rel_err_vs_fp64": 1.9e-6},
                                   "forward_plus_adjoint": {"speedup_vs_fp64_gpu": 61.0}},
                        "split2": {"rel_err_vs_fp64": 3.6e-6,
                                   "adjoint": {"rel_err_vs_fp64": 6.1e-6},
                                   "forward_plus_adjoint": {"speedup_vs_fp64_gpu": 102.0}},
                        "fp16x2": {"rel_err_vs_fp64": 1.5e-6,
                                   "adjoint": {"rel_err_vs_fp64": 1.7e-6},
                                   "forward_plus_adjoint": {
                                       "speedup_vs_fp64_gpu": 94.0,
                                       "speedup_vs_fp64_same_algorithm": 7.0}}}}
    s = b.summary(line)
    assert s["fwd_adj_vs_fp64_gpu"] == {"point_mass_fp32": 3.2, "prism_fp32": 220.0,
                                        "prism_fp32_vs_fp64_same_algorithm": 3.1,
                                        "ensemble_split3": 61.0, "ensemble_split2": 102.0,
                                        "ensemble_fp16x2": 94.0,
                                        "ensemble_fp16x2_vs_fp64_same_algorithm": 7.0}
    assert s["point_mass_vs_cpu_baseline"] == 4800.0
    assert s["gn_step_vs_cpu_baseline"] is None
    assert s["max_rel_err_vs_fp64"] == {"point_mass_fp32": 3e-7, "prism_fp32": 5e-7,
                                        "ensemble_split3": 1.9e-6, "ensemble_split2": 6.1e-6,
                                        "ensemble_fp16x2": 1.7e-6}


def test_summary_without_extra_legs():
    b = load_bench()
    s = b.summary({"speedup_vs_fp64_gpu": 3.0, "rel_err_vs_fp64": {"forward": 1e-7}})
    assert s["fwd_adj_vs_fp64_gpu"]["ensemble_split3"] is None
    assert s["max_rel_err_vs_fp64"]["ensemble_split2"] is None
    assert s["max_rel_err_vs_fp64"]["ensemble_fp16x2"] is None
    assert s["point_mass_vs_cpu_baseline"] is None
    assert s["gn_step_normal_equations_mfma_frac"] is None
    s = b.summary({"gn_step": {"speedup_vs_cpu": 4000.0, "roofline": {"frac": 0.5}}})
    assert s["gn_step_vs_cpu_baseline"] == 4000.0
    assert s["gn_step_normal_equations_mfma_frac"] == 0.5


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=300)
    assert r.returncode != 0
    assert "one GPU" in (r.stdout + r.stderr)
