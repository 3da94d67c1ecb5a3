"""Run-time tuning of the plugin (prgpu_tune_set / prgpu_tune_get / prgpu_tune_source), on the CPU.

The tuning file (written by ar-renderer_amd/autotune.py) is read when the plugin first
enumerates devices; that needs no GPU, so the parsing is checked here in fresh processes: the
file named by PRGPU_TUNING_FILE is found, a valid file reports no error, and the first invalid line
(unknown key, value out of range, odd graph_rounds, trailing text) is named.
"""
import os
import subprocess
import sys
import textwrap

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

PROBE = textwrap.dedent("""
    import ctypes as C, sys
    sys.path.insert(0, {pkg!r})
    import prender_amd as pa
    p = pa.Plugin()
    t = p.tuning()
    print(t["file"] or "")
    print(t["file_error"] or "")
""")


def _probe(tmp_path, text):
    f = tmp_path / "tuning.txt"
    f.write_text(text)
    env = dict(os.environ, PRGPU_TUNING_FILE=str(f))
    out = subprocess.run([sys.executable, "-c", PROBE.format(pkg=os.path.join(ROOT, "ar-renderer_amd"))], env=env,
                         capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    path, err = (out.stdout.split("\n") + ["", ""])[:2]
    return str(f), path, err


def test_valid_tuning_file_is_read(tmp_path):
    f, path, err = _probe(tmp_path, "# written by the autotuner\ntrace_blocks_per_cu 6\nshade_blocks_per_cu 12  # comment\n"
                                    "\npath_budget 4194304\nguide_lr 0.003\ngraph_rounds 4\nmaterial_sort 1\nray_sort 1\n")
    assert path == f
    assert err == ""


def test_invalid_tuning_lines_are_named(tmp_path):
    for bad in ("no_such_key 3", "trace_blocks_per_cu 9", "graph_rounds 3", "guide_batch 1000.5", "guide_lr 0.1 extra",
                "shade_blocks_per_cu", "material_sort 2", "ray_sort -1"):
        f, path, err = _probe(tmp_path, "trace_blocks_per_cu 4\n" + bad + "\n")
        assert path == f
        assert err.endswith(":2: invalid tuning line"), (bad, err)


def test_missing_tuning_file_means_defaults(tmp_path):
    env = dict(os.environ, PRGPU_TUNING_FILE=str(tmp_path / "absent.txt"))
    out = subprocess.run([sys.executable, "-c", PROBE.format(pkg=os.path.join(ROOT, "ar-renderer_amd"))], env=env,
                         capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    assert out.stdout.split("\n")[:2] == ["", ""]


def test_committed_tuning_file_beside_the_library_is_valid():
    """The autotuner's file shipped beside libprender_gpu.so is what the plugin reads by default."""
    env = {k: v for k, v in os.environ.items() if k != "PRGPU_TUNING_FILE"}
    out = subprocess.run([sys.executable, "-c", PROBE.format(pkg=os.path.join(ROOT, "ar-renderer_amd"))], env=env,
                         capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    path, err = (out.stdout.split("\n") + ["", ""])[:2]
    assert os.path.realpath(path) == os.path.realpath(os.path.join(ROOT, "ar-renderer_amd", "prgpu_tuning.txt"))
    assert err == ""
