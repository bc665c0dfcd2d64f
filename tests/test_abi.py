"""The C-ABI library: it loads without a GPU, exports every symbol the headers
declare, and its structs have the reference layout (no compute calls here)."""
import ctypes as C
import os
import re
import subprocess

import pytest

from corrfunc_amd import _abi, _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
INC = os.path.join(ROOT, "include")


def declared_functions():
    names = []
    for h in ("corrfunc_theory.h", "corrfunc_amd.h", "corrfunc_io.h", "corrfunc_mocks.h"):
        text = open(os.path.join(INC, h)).read()
        text = re.sub(r"static inline[^{]*\{(?:[^{}]|\{[^{}]*\})*\}", "", text)
        names += re.findall(r"^\s*(?:int64_t|int|void|const char\s*\*)\s+\**\s*(\w+)\s*\(", text, re.M)
    return sorted(set(names))


def test_library_exports_every_declared_symbol():
    lib = C.CDLL(_lib.LIB_PATH)
    names = declared_functions()
    assert {"countpairs", "countpairs_rp_pi", "countpairs_s_mu", "countpairs_xi", "countpairs_wp",
            "free_results", "free_results_rp_pi", "free_results_s_mu", "free_results_xi",
            "free_results_wp", "cf_amd_last_stats", "read_positions", "read_columns_into_array", "countpairs_mocks", "countpairs_mocks_s_mu",
            "free_results_mocks", "free_results_mocks_s_mu"} <= set(names)
    for n in names:
        assert hasattr(lib, n), n
    nm = subprocess.run(["nm", "-D", "--defined-only", _lib.LIB_PATH], capture_output=True, text=True).stdout
    for n in names:
        assert re.search(r"\sT\s%s$" % n, nm, re.M), n


def test_header_layout_matches_ctypes(tmp_path):
    src = tmp_path / "probe.c"
    src.write_text(r'''
#include <stdio.h>
#include <stddef.h>
#include "corrfunc_theory.h"
#include "corrfunc_amd.h"
int main(void) {
  struct config_options o = get_config_options();
  struct extra_options e = get_extra_options(PAIR_PRODUCT);
  printf("%zu %zu %zu %zu %zu %zu %zu\n", sizeof(struct config_options), sizeof(struct extra_options),
         offsetof(struct config_options, instruction_set), offsetof(struct config_options, bin_refine_factors),
         offsetof(struct config_options, binning_flags), offsetof(struct extra_options, weight_method),
         sizeof(cf_amd_stats));
  printf("%s %d %d %d %d %d %ld\n", o.version, o.bin_refine_factors[0], o.bin_refine_factors[1],
         o.bin_refine_factors[2], o.max_cells_per_dim, o.periodic, (long)e.weights0.num_weights);
  return 0;
}''')
    exe = tmp_path / "probe"
    subprocess.check_call(["gcc", "-std=c99", "-Wall", "-Werror", "-I", INC, str(src), "-o", str(exe)])
    out = subprocess.check_output([str(exe)], text=True).split("\n")
    sizes = [int(v) for v in out[0].split()]
    assert sizes == [1024, 1024, _abi.config_options.instruction_set.offset,
                     _abi.config_options.bin_refine_factors.offset, _abi.config_options.binning_flags.offset,
                     _abi.extra_options.weight_method.offset, C.sizeof(_lib.cf_amd_stats)]
    assert out[1].split() == ['"2.5.3"', "2", "2", "1", "100", "1", "1"]


def test_header_matches_reference_layout(tmp_path):
    """Same offsets as the reference's utils/defs.h, when the reference is mounted."""
    ref_utils = "/root/reference/utils"
    if not os.path.isdir(ref_utils):
        pytest.skip("reference sources not mounted")
    fields = ["boxsize_x", "c_api_time", "cell_timings", "totncells_timings", "float_type", "instruction_set",
              "version", "verbose", "c_api_timer", "c_cell_timer", "need_avg_sep", "autocorr", "periodic",
              "sort_on_z", "fast_divide_and_NR_steps", "enable_min_sep_opt", "bin_refine_factors",
              "max_cells_per_dim", "copy_particles", "use_heap_sort", "binning_flags"]
    body = "".join('printf("%%zu ", offsetof(struct config_options, %s));' % f for f in fields)
    body += 'printf("%zu %zu\\n", offsetof(struct extra_options, weights1), offsetof(struct extra_options, weight_method));'
    outs = []
    for inc, extra in ((INC, '#include "corrfunc_theory.h"'), (ref_utils, '#include "defs.h"')):
        src = tmp_path / ("p%d.c" % len(outs))
        src.write_text("#include <stdio.h>\n#include <stddef.h>\n%s\nint main(void){%s return 0;}\n" % (extra, body))
        exe = tmp_path / ("p%d" % len(outs))
        subprocess.check_call(["gcc", "-std=gnu99", "-w", '-DVERSION="2.5.3"', "-I", inc, str(src), "-o", str(exe)])
        outs.append(subprocess.check_output([str(exe)], text=True))
    assert outs[0] == outs[1]
