// Drop-in public API (reference include/interface.h:6-10): one call per correlator family.
// The perambulator is the host buffer returned by contract::readPropH (layout
// [Nt][Nt][4][4][NeV][NeV] of Float), the elementals are FP64 driver arrays [type][Nt][NeV][NeV].
#ifndef INTERFACE_H
#define INTERFACE_H

#include <contract.h>

void calculateRho(contract::contractInfo, void*, double*, contract::PRECISION, std::string outfile);
void calculateRhoPi(int idir, contract::contractInfo, void*, double*, double*, double*,
                    contract::PRECISION, std::string outfile);
void calculatePiPi(contract::contractInfo, void*, double*, double*, double*, double*,
                   contract::PRECISION, std::string outfile, contract::ISOSPIN);
void calculateSigma(contract::contractInfo, void*, double*, contract::PRECISION, std::string outfile);
void calculateSigmaPi(contract::contractInfo, void*, double*, double*, double*, contract::PRECISION,
                      std::string outfile);
#endif
