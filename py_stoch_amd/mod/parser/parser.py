"""Shared parsing helpers (reference: mod/parser/parser.py:7-71)."""
import xml.etree.ElementTree as ET

import numpy as np


class Parser:
    @staticmethod
    def parse_settings(settings_file, args_obj):
        """Read output species, duration, sample count and simulation count
        from a settings XML file (parser.py:9-29)."""
        root = ET.parse(settings_file).getroot()
        out_species = root.find("output_species")
        args_obj.E = np.array([int(s.get("idx")) for s in out_species.findall("species")]).astype(np.int32)
        args_obj.t_max = float(root.find("duration").text)
        args_obj.ita = int(root.find("num_output_time_points").text)
        args_obj.U = int(root.find("num_simulations").text)

    @staticmethod
    def define_hazards(spn):
        """C source of UpdatePropensities as the reference generates it
        (parser.py:32-55).  Kept for API compatibility; the HIP kernels use
        the mass-action table instead of compiled hazard code."""
        lines = ""
        for idx, haz in enumerate(spn.h):
            lines += "    a[" + str(idx) + "] = " + haz + ";\n"
        code = ("\n            __device__ void UpdatePropensities(double* __restrict__ a, "
                "int* __restrict__ x)\n            {\n" + lines + "\n            }\n            ")
        for i in range(len(spn.P), 0, -1):
            code = code.replace("species_" + str(i - 1), "x[" + str(i - 1) + "]")
        for i in range(len(spn.c), 0, -1):
            code = code.replace("parameter_" + str(i - 1), "d_c[" + str(i - 1) + "]")
        return code

    @staticmethod
    def flatten_matrix(matrix):
        """Non-zeros of a species x reaction matrix as (species, reaction,
        value, 0) rows in species-major order (parser.py:58-71).

        The reference packs them as int8 char4 with a ubyte count (at most
        127 species/reactions and 255 non-zeros); here they are int32 rows
        and an int count."""
        matrix = np.asarray(matrix)
        rows = []
        for i in range(matrix.shape[0]):
            for j in range(matrix.shape[1]):
                if matrix[i][j] != 0:
                    rows.append([i, j, int(matrix[i][j]), 0])
        return np.array(rows, dtype=np.int32).reshape(-1, 4), len(rows)
