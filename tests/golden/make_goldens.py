"""Extracts the reference's regression values into tests/golden/runtest_goldens.json
and copies the BC option files used by the tests.

Run in a container where the reference tree is mounted (default /root/reference):
    python tests/golden/make_goldens.py [REFERENCE_ROOT]
Only data is taken: the printed energies / crack openings of
ValidationTests/results/runtest1.out and runtest2.out (the two goldens that
match the current drivers) and the BCfiles/*.txt option lists.
"""
import json
import os
import re
import shutil
import sys

ref = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
here = os.path.dirname(os.path.abspath(__file__))

steps = []
cur = None
for line in open(os.path.join(ref, "ValidationTests/results/runtest1.out")):
    m = re.match(r"==== time step (\d+): t = (\S+)", line)
    if m:
        cur = dict(step=int(m.group(1)), t=float(m.group(2)), altmin=0)
        steps.append(cur)
        continue
    if cur is None:
        continue
    if "Max. change on V" in line:
        cur["altmin"] += 1
    m = re.match(r"Total crack opening: (\S+)", line)
    if m:
        cur["crack_opening"] = float(m.group(1))
    m = re.match(r"Elastic Energy:\s+(\S+)", line)
    if m:
        cur["elastic"] = float(m.group(1))
    m = re.match(r"Surface  energy:\s+(\S+)", line)
    if m:
        cur["surface"] = float(m.group(1))

t2 = {}
for line in open(os.path.join(ref, "ValidationTests/results/runtest2.out")):
    for key, pat in (("surface", r"Surface energy:\s+(\S+)"), ("elastic", r"Elastic Energy:\s+(\S+)"),
                     ("pressure_work", r"Work of pressure forces:\s+(\S+)")):
        m = re.match(pat, line)
        if m:
            t2[key] = float(m.group(1))

out = dict(
    runtest1=dict(source="ValidationTests/results/runtest1.out",
                  args="-n 101,5,5 -l 10,1,1 -nu 0 -epsilon 0.5 -eta 1e-8 -gc .1 -altmintol 1.e-3 "
                       "-maxtimestep 11 -maxtimevalue 4 -atnum 1 -V_X0_BC ONE -U_X0_BC_0 FIXED -U_X0_0 -1. "
                       "-U_X0_BC_1 ZERO -U_X0_BC_2 ZERO -V_X1_BC ONE -U_X1_BC_0 FIXED -U_X1_0 1. "
                       "-U_X1_BC_1 ZERO -U_X1_BC_2 ZERO -verbose 1 -alpha 0 -beta 0 -pressure 0",
                  steps=steps),
    runtest2=dict(source="ValidationTests/results/runtest2.out",
                  args="-n 41,81,6 -l 1,2,.25 -length .4 -orientation 3 -nu 0 -epsilon 0.05 -eta 1e-8 "
                       "-altmintol 1.e-3", **t2),
)
json.dump(out, open(os.path.join(here, "runtest_goldens.json"), "w"), indent=1)
for f in ("3DClamped.txt", "3DRollers.txt", "3DSYMXYZ.txt", "3DRigidMotion.txt"):
    shutil.copy(os.path.join(ref, "BCfiles", f), os.path.join(here, "bcfiles", f))
print("wrote", os.path.join(here, "runtest_goldens.json"))
