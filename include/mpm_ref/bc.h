// Reference-signature host shim: node-list Dirichlet boundary conditions.
//
// BoundaryCondition with the members and methods of the reference's
// module/bc.h:7-22 (definitions module/bc.cpp:9-79): BCInput / BCOutput read
// and write "ibc ! ---- name.ibc ----" followed by ibc lines "nbc fbc";
// BCSetVal sets variable[nbc + nn] = fbc * facl, BCSetZero sets it to zero,
// BCSetDt to fbc * dt * facl (facl and dt are the globals of dataset.h).
#pragma once

#include <fstream>
#include <iostream>
#include <string>
#include <vector>

class BoundaryCondition {
  public:
    int ibc = 0;
    std::vector<int> nbc;
    std::vector<double> fbc;

    void BCInput(std::ifstream &infile, bool if_fbc = true);

    void BCOutput(std::ofstream &outfile, std::string bc_name, bool ife_fbc = true);

    void BCSetVal(int nn, std::vector<double> &variable);

    void BCSetZero(int nn, std::vector<double> &variable);

    void BCSetDt(int nn, std::vector<double> &variable);
};
