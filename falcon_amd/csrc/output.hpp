// Host side of the per-block AllCells output: the reference's text format
// (ModSave_AllCells, 3_6/ModSavePlot.f90:520-650) and its merged Tecplot form
// (3_6/merge_AllCells.py).  The corner values themselves are formed on the
// device (k_allcells); this file only formats and writes them.
#pragma once
#include <cstdint>
#include <string>
#include <vector>

#include "layout.hpp"

namespace falcon {

// Fortran ES16.8 edit of one value into out[0..15] (no terminator): right-
// justified d.dddddddd with a signed two-digit exponent "E+dd", or "+ddd"
// without the E where the exponent needs three digits (Fortran's ESw.d form).
void format_es16(double x, char *out);

struct AllCellsSpec {
    int ni, nj, nk, nvar_total;   // nvar_total: 3 coordinates + the reference's nvar (9)
    int nblocks;
    std::string variables;        // the VARIABLES line of info.dat
};

// info.dat of one output directory, as rank 0 of ModSave_AllCells writes it.
std::string allcells_info(const AllCellsSpec &s);

// The text of one block: (ni+1)(nj+1)(nk+1) lines of nvar_total ES16.8 fields,
// from vals[corner][nvar_total] in the reference's loop order (i fastest).
void allcells_block_text(const AllCellsSpec &s, const double *vals, std::string &out);

// Per block: face radii xi_F(1:ni+1) as ModBlock_InitGrid forms them
// (3_6/ModBlock.f90:320-322), then sin and cos of the theta faces and of the
// phi faces (libm, as the reference's writer calls them); and the Yin flag.
void allcells_geometry(const Layout &lay, std::vector<double> &geo, std::vector<int32_t> &yin);

// Connectivity lines of merge_AllCells.py for blocks [b0, b1).
void allcells_connectivity(const AllCellsSpec &s, int b0, int b1, std::string &out);

}  // namespace falcon
