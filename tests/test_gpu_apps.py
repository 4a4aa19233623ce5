"""The bfs / pr apps (the reference's command line, src/bfs.cc:238-278,
src/pr.cc:120-150) on a graph given as a .sg file (-f, src/builder.h:320-322):
each trial's result passes the app's verifier (gapbs-verify/bfs.cc:195-241,
gapbs-verify/pr.cc:77-93)."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "dist_mem_gapbs_amd", "bin")


@pytest.mark.parametrize("app", ["bfs", "pr"])
def test_app_reads_sg_and_verifies(dmm, graph, tmp_path, app):
    exe = os.path.join(BIN, app)
    if not os.path.exists(exe):
        pytest.skip("apps not built")
    g = graph(12, 16, False)
    p = tmp_path / "kron12.sg"
    g.write_serialized(p)
    print(f"[app {app} -f {p.name}]", flush=True)
    r = subprocess.run([exe, "-f", str(p), "-n", "2", "-v"], capture_output=True, text=True, timeout=100,
                       cwd=tmp_path)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    lines = [ln for ln in r.stdout.splitlines() if "Verification" in ln]
    assert len(lines) == 2 and all("PASS" in ln for ln in lines), r.stdout[-2000:]
    assert "Read Time" in r.stdout


def test_bfs_app_step_lines_match_between_loops(graph, tmp_path):
    """The bfs app runs the device-resident loop by default and the
    host-driven search with DMM_STEP_TIMES=1: both print the same step
    lines (kind and count, src/bfs.cc:149-184), and both verify."""
    exe = os.path.join(BIN, "bfs")
    if not os.path.exists(exe):
        pytest.skip("apps not built")
    g = graph(12, 16, False)
    p = tmp_path / "kron12.sg"
    g.write_serialized(p)

    def steps(env_extra):
        env = dict(os.environ, **env_extra)
        r = subprocess.run([exe, "-f", str(p), "-n", "3", "-v"], capture_output=True, text=True, timeout=100,
                           cwd=tmp_path, env=env)
        assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
        assert r.stdout.count("PASS") == 3, r.stdout[-2000:]
        out = []
        for ln in r.stdout.splitlines():
            # "PE: 0 |    td         12     0.00012" (harness.h PrintStep)
            parts = ln.split("|", 1)[1].split() if ln.startswith("PE:") and "|" in ln else []
            if parts and parts[0] in ("td", "bu", "e", "c"):
                out.append((parts[0], parts[1] if parts[0] in ("td", "bu") else None))
        return out

    loop, host = steps({"DMM_STEP_TIMES": "0"}), steps({"DMM_STEP_TIMES": "1"})
    assert loop and loop == host
