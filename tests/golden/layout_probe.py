"""Struct layout probe: compiles a small C program against a header that defines the
CTDSE structs and returns {struct: {"size": n, "fields": {name: offset}}}."""
import json
import os
import subprocess
import tempfile

FIELDS = {
    "Efield_var": ["tgrid", "Field", "dt", "omega", "E0", "phi", "ton", "toff", "Nt", "nc"],
    "trg_def": ["a"],
    "analy_def": ["tprint", "writewft"],
    "output_print_def": ["Efield", "FEfield", "sourceterm", "Fsourceterm", "FEfieldM2", "FsourceTermM2", "PopTot",
                         "tgrid", "omegagrid", "PopInt", "expval_x"],
    "inputs_def": ["trg", "Efield", "Eguess", "Einit", "tmin", "Nt", "num_t", "dt", "num_r", "dx", "psi0", "x",
                   "gauge", "x_int", "analy", "InterpByDTorNT", "Ntinterp", "Print", "CV", "precision"],
    "outputs_def": ["tgrid", "Efield", "sourceterm", "omegagrid", "FEfield", "Fsourceterm", "FEfieldM2",
                    "FsourcetermM2", "PopTot", "PopInt", "expval", "Nt", "Nomega", "psi"],
}


HANKEL_FIELDS = {
    "hk_medium": ["mode", "pressure", "np_z", "np_r", "p_z", "p_r", "p_value", "include_dispersion",
                  "include_absorption", "n_ir", "n_f1", "n_f2", "e_f1", "f1", "e_f2", "f2"],
    "hk_planes": ["nz", "no", "nr", "nrho", "src", "src_c64", "plane_stride", "zgrid", "ogrid", "rgrid", "tw", "rho",
                  "distance", "near_field", "pre", "pre_mode"],
}


def probe(include_dir, header, structs=None):
    table = dict(FIELDS, **HANKEL_FIELDS)
    names = list(structs) if structs else list(FIELDS)
    lines = ["#include <stdio.h>", "#include <stddef.h>", '#include "%s"' % header, "int main(void) {",
             'printf("{");']
    first = True
    for s in names:
        fields = table[s]
        lines.append('printf("%s\\"%s\\": {\\"size\\": %%zu, \\"fields\\": {", sizeof(%s));' % ("" if first else ",", s, s))
        first = False
        for i, f in enumerate(fields):
            lines.append('printf("%s\\"%s\\": %%zu", offsetof(%s, %s));' % ("" if i == 0 else ",", f, s, f))
        lines.append('printf("}}");')
    lines += ['printf("}\\n");', "return 0;", "}"]
    with tempfile.TemporaryDirectory() as d:
        c = os.path.join(d, "probe.c")
        exe = os.path.join(d, "probe")
        open(c, "w").write("\n".join(lines))
        subprocess.run(["gcc", "-w", "-I", include_dir, c, "-o", exe], check=True)
        return json.loads(subprocess.run([exe], check=True, capture_output=True, text=True).stdout)
