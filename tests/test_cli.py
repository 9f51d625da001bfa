"""The imm command-line tool (ripples_amd/tools/imm.cpp): the reference's flags
(tools/imm.cc; App. A of the flag list in include/ripples/configuration.h)
and its JSON experiment record (tools/imm.cc:101-136)."""
import json
import os
import subprocess

import numpy as np
import pytest

import ripples_amd as R
from graphs import GOLDEN

IMM = os.path.join(os.path.dirname(R.__file__), "bin", "imm")
SIM = os.path.join(os.path.dirname(R.__file__), "bin", "simulator")
KEYS = ["Algorithm", "Input", "Output", "DiffusionModel", "Epsilon", "K", "L", "NumThreads",
        "NumWalkWorkers", "NumCPUTeams", "NumGPUWalkWorkers", "PauseThreshold", "Total",
        "ThetaPrimeDeltas", "ThetaEstimation", "ThetaEstimationGenerateRRR",
        "ThetaEstimationMostInfluential", "Theta", "Counting", "Pivoting",
        "Microbenchmarking", "CPUBatchSize", "GPUBatchSize", "RRRSetSizeBytes",
        "GenerateRRRSets", "FindMostInfluentialSet", "Seeds"]


def run(*args, tool=None):
    return subprocess.run([tool or IMM, *args], capture_output=True, text=True, timeout=300)


@pytest.fixture(scope="module", autouse=True)
def built():
    if not (os.path.exists(IMM) and os.path.exists(SIM)):
        pytest.skip("command-line tools not built")


@pytest.mark.parametrize("args,msg", [
    ([], "--input-graph is required"),
    (["-i", "g.tsv", "-e", "0.5", "-d", "IC"], "--seed-set-size is required"),
    (["-i", "g.tsv", "-k", "5", "-d", "IC"], "--epsilon is required"),
    (["-i", "g.tsv", "-k", "5", "-e", "0.5"], "--diffusion-model is required"),
    (["-i", "g.tsv", "-k", "5", "-e", "0.5", "-d", "XX"], "must be IC or LT"),
    (["-i", "g.tsv", "-k", "5", "-e", "0.5", "-d", "IC", "--bogus"], "unknown option"),
])
def test_usage_errors(args, msg):
    r = run(*args)
    assert r.returncode == 2 and msg in r.stderr


def test_missing_input_file(tmp_path):
    r = run("-i", str(tmp_path / "none.tsv"), "-k", "3", "-e", "0.5", "-d", "IC")
    assert r.returncode == 1 and "cannot open" in r.stderr


@pytest.mark.gpu
@pytest.mark.parametrize("model,extra", [("IC", ["-w"]), ("LT", [])])
def test_imm_record_matches_library(tmp_path, model, extra):
    out = str(tmp_path / "out.json")
    r = run("-i", GOLDEN + "/karate.tsv", "-k", "4", "-e", "0.3", "-d", model, "-p",
            "-o", out, *extra)
    assert r.returncode == 0, r.stderr
    rec = json.load(open(out))
    assert isinstance(rec, list) and len(rec) == 1
    rec = rec[0]
    assert list(rec) == KEYS
    assert rec["DiffusionModel"] == model and rec["K"] == 4 and rec["Epsilon"] == 0.3
    # the same run through the library: seeds in original ids
    mdl = R.IC if model == "IC" else R.LT
    hg = R.load_edge_list(GOLDEN + "/karate.tsv", model=mdl, weighted=bool(extra))
    g = R.Graph(hg.index, hg.edges, hg.weights, model=mdl)
    seeds, lib_rec = R.imm(g, 4, 0.3, seed=0)
    assert rec["Seeds"] == hg.reverse_map[seeds].tolist()
    assert rec["Theta"] == lib_rec["Theta"]
    assert sum(rec["ThetaPrimeDeltas"]) == max(lib_rec["Theta"], max(lib_rec["ThetaPrime"]))
    g.close()


@pytest.mark.gpu
def test_rrsets_mode_and_binary_reload(tmp_path):
    hg = R.load_edge_list(GOLDEN + "/karate.tsv", weighted=True)
    binp = str(tmp_path / "k.bin")
    hg.dump_binary(binp)
    out = str(tmp_path / "r.json")
    r = run("-i", binp, "--reload-binary", "-k", "3", "-e", "0.5", "-d", "IC",
            "--rrsets", "5000", "-o", out)
    assert r.returncode == 0, r.stderr
    rec = json.load(open(out))[0]
    assert rec["Theta"] == 5000 and rec["ThetaPrimeDeltas"] == [5000]
    assert rec["Seeds"] == [int(hg.reverse_map[1])] * 3
    assert rec["RRRSetSizeBytes"] > 4 * 5000


@pytest.mark.parametrize("args,msg", [
    ([], "--input-graph is required"),
    (["-i", "g.tsv", "--replicas", "5"], "--experiment-file is required"),
    (["-i", "g.tsv", "-e", "x.json"], "--replicas is required"),
])
def test_simulator_usage(args, msg):
    r = run(*args, tool=SIM)
    assert r.returncode == 2 and msg in r.stderr


def test_simulator_bad_record(tmp_path):
    bad = tmp_path / "bad.json"
    bad.write_text('[{"DiffusionModel": "XX", "Seeds": [1]}]')
    r = run("-i", GOLDEN + "/karate.tsv", "-e", str(bad), "--replicas", "3", tool=SIM)
    assert r.returncode == 1 and "IC or LT" in r.stderr
    bad.write_text('[{"DiffusionModel": "IC", "Seeds": [1,]')
    r = run("-i", GOLDEN + "/karate.tsv", "-e", str(bad), "--replicas", "3", tool=SIM)
    assert r.returncode == 1 and "bad experiment file" in r.stderr


@pytest.mark.gpu
@pytest.mark.parametrize("model,extra", [("IC", ["-w"]), ("LT", [])])
def test_imm_then_simulator(tmp_path, model, extra, orc):
    """imm -> simulator round trip; the simulations equal the oracle's spread
    of the same seeds (tools/simulator.cc:100-171)."""
    rec_path, out = str(tmp_path / "imm.json"), str(tmp_path / "sim.json")
    r = run("-i", GOLDEN + "/karate.tsv", "-k", "3", "-e", "0.5", "-d", model, "-o", rec_path,
            *extra)
    assert r.returncode == 0, r.stderr
    r = run("-i", GOLDEN + "/karate.tsv", "-e", rec_path, "--replicas", "300", "-o", out, *extra,
            tool=SIM)
    assert r.returncode == 0, r.stderr
    imm_rec = json.load(open(rec_path))[0]
    sim = json.load(open(out))
    assert len(sim) == 1 and list(sim[0]) == ["Input", "Algorithm", "DiffusionModel", "Epsilon",
                                               "K", "Seeds", "Simulations"]
    assert sim[0]["Seeds"] == imm_rec["Seeds"] and sim[0]["DiffusionModel"] == model
    mdl = R.IC if model == "IC" else R.LT
    hg = R.load_edge_list(GOLDEN + "/karate.tsv", model=mdl, weighted=bool(extra))
    internal = [int(np.searchsorted(hg.reverse_map, s)) for s in imm_rec["Seeds"]]
    act, lev = orc.spread(mdl, hg.index, hg.edges, hg.weights, internal, 0, 0, 300)
    assert sim[0]["Simulations"] == [[int(a), int(b)] for a, b in zip(act, lev)]


OPIMC = os.path.join(os.path.dirname(R.__file__), "bin", "opimc")
OPIMC_KEYS = ["Algorithm", "Input", "Output", "DiffusionModel", "Epsilon", "K", "Delta",
              "NumThreads", "NumWalkWorkers", "NumCPUTeams", "NumGPUWalkWorkers",
              "PauseThreshold", "Total", "CPUBatchSize", "GPUBatchSize", "RRRSetSizeBytes",
              "GenerateRRRSets", "FindMostInfluentialSet", "Microbenchmarking", "ThetaZero",
              "ThetaMax", "RRRSetsGenerated", "Seeds"]


@pytest.mark.parametrize("args,msg", [
    (["-i", "g.tsv", "-k", "5", "-e", "0.5"], "--diffusion-model is required"),
    (["-i", "g.tsv", "-k", "5", "-e", "0.5", "-d", "IC", "--rrsets", "9"], "unknown option"),
])
def test_opimc_usage_errors(args, msg):
    r = run(*args, tool=OPIMC)
    assert r.returncode == 2 and msg in r.stderr


@pytest.mark.gpu
@pytest.mark.parametrize("model,extra", [("IC", ["-w"]), ("LT", [])])
def test_opimc_record_matches_library(tmp_path, model, extra):
    """tools/opimc.cc:63-97 record; the seeds are rip_opimc's in original ids."""
    out = str(tmp_path / "o.json")
    r = run("-i", GOLDEN + "/karate.tsv", "-k", "3", "-e", "0.3", "-d", model, "--delta", "0.1",
            "-o", out, *extra, tool=OPIMC)
    assert r.returncode == 0, r.stderr
    rec = json.load(open(out))[0]
    assert list(rec) == OPIMC_KEYS and rec["Algorithm"] == "OPIMC"
    mdl = R.IC if model == "IC" else R.LT
    hg = R.load_edge_list(GOLDEN + "/karate.tsv", model=mdl, weighted=bool(extra))
    g = R.Graph(hg.index, hg.edges, hg.weights, model=mdl)
    seeds, lib_rec = R.opimc(g, 3, 0.3, 0.1, seed=0)
    assert rec["Seeds"] == hg.reverse_map[seeds].tolist()
    assert rec["ThetaZero"] == lib_rec["ThetaZero"] and rec["ThetaMax"] == lib_rec["ThetaMax"]
    assert rec["RRRSetsGenerated"] == lib_rec["RRRSetsGenerated"]
    assert abs(rec["Delta"] - 0.1) < 1e-7
    g.close()
