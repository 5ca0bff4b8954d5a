"""The ctypes mirrors of the C structs have the C layout (size and offsets),
checked against a tiny C program compiled from include/jvmc_hip.h."""
import ctypes
import os
import subprocess
import tempfile

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

SRC = r"""
#include <stddef.h>
#include <stdio.h>
#include "jvmc_hip.h"
#define P(T, f) printf(#T "." #f " %zu\n", offsetof(T, f))
int main(void) {
  printf("jv_net %zu\njv_optable %zu\njv_mc_args %zu\njv_solve_info %zu\njv_orbit %zu\njv_batch_item %zu\n",
         sizeof(jv_net), sizeof(jv_optable), sizeof(jv_mc_args), sizeof(jv_solve_info), sizeof(jv_orbit),
         sizeof(jv_batch_item));
  P(jv_batch_item, params); P(jv_batch_item, op); P(jv_batch_item, prefactor); P(jv_batch_item, mc);
  P(jv_batch_item, walkers); P(jv_batch_item, grad); P(jv_batch_item, grad_layout);
  P(jv_net, precision); P(jv_net, num_sites); P(jv_net, num_params); P(jv_net, off_conv_kernel); P(jv_net, off_conv_bias);
  P(jv_optable, max_sites_changed); P(jv_optable, idx); P(jv_optable, diag); P(jv_optable, fermi); P(jv_mc_args, mu); P(jv_mc_args, seed); P(jv_mc_args, stream);
  P(jv_solve_info, rank); P(jv_solve_info, e_var); P(jv_solve_info, tdvp_error); P(jv_solve_info, e_mean);
  P(jv_orbit, avg); P(jv_orbit, complex_factor); P(jv_orbit, map); P(jv_orbit, factor);
  return 0;
}
"""


def test_ctypes_struct_layouts_match_c():
    from vmc_jax_amd import _lib as L
    with tempfile.TemporaryDirectory() as d:
        c = os.path.join(d, "s.c")
        exe = os.path.join(d, "s")
        open(c, "w").write(SRC)
        subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), c, "-o", exe], check=True)
        out = subprocess.run([exe], capture_output=True, text=True, check=True).stdout.split("\n")
    vals = dict(l.split() for l in out if l.strip())
    assert int(vals["jv_net"]) == ctypes.sizeof(L.JvNet)
    assert int(vals["jv_optable"]) == ctypes.sizeof(L.JvOpTable)
    assert int(vals["jv_mc_args"]) == ctypes.sizeof(L.JvMcArgs)
    assert int(vals["jv_solve_info"]) == ctypes.sizeof(L.JvSolveInfo)
    assert int(vals["jv_orbit"]) == ctypes.sizeof(L.JvOrbit)
    assert int(vals["jv_batch_item"]) == ctypes.sizeof(L.JvBatchItem)
    m = {"jv_net": L.JvNet, "jv_optable": L.JvOpTable, "jv_mc_args": L.JvMcArgs, "jv_solve_info": L.JvSolveInfo,
         "jv_orbit": L.JvOrbit, "jv_batch_item": L.JvBatchItem}
    for k, v in vals.items():
        if "." in k:
            t, f = k.split(".")
            assert getattr(m[t], f).offset == int(v), k
