"""Writes tests/golden/dadi_sfs.json from the reference's committed outputs of
its dadi validation scenario (examples/example_dadi/):
  run_980.out    L = 2e9: per-bin mean, SD, CV of the normalised n = 1001 SFS
                 of population 1 over 50 runs (GTX 980)
  run_980_2.out  the same at L = 2e8
  dadi_example.out  the diffusion (dadi) solution, bins 1..1000
Run with /root/reference mounted.
"""
import json
import os

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = "/root/reference/examples/example_dadi"


def read_run(name):
    mean, sd = [], []
    tail = {}
    with open(os.path.join(SRC, name)) as f:
        for line in f:
            t = line.strip().split("\t")
            if len(t) == 4 and t[0].isdigit():
                mean.append(float(t[1]))
                sd.append(float(t[2]))
            elif ":" in line:
                k, v = line.split(":", 1)
                try:
                    tail[k.strip()] = float(v)
                except ValueError:
                    pass
    return dict(mean=mean, sd=sd, **{k.replace(" ", "_"): v for k, v in tail.items()})


def read_dadi():
    vals = []
    with open(os.path.join(SRC, "dadi_example.out")) as f:
        lines = f.read().split("\n")
    for line in lines:
        try:
            vals.append(float(line))
        except ValueError:
            continue
    return vals


def main():
    data = {"run_980": read_run("run_980.out"), "run_980_2": read_run("run_980_2.out"),
            "dadi": read_dadi(), "source": "reference examples/example_dadi/*.out"}
    with open(os.path.join(HERE, "dadi_sfs.json"), "w") as f:
        json.dump(data, f)


if __name__ == "__main__":
    main()
