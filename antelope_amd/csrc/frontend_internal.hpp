// frontend_internal.hpp - the parsed-input object shared by the config-file
// readers (frontend.cpp: inputParam.cfg, atl.cpp: .atl). Not part of the ABI.
#pragma once

#include <string>
#include <vector>

#include "antelope/sbe_frontend.h"

struct sbe_spec {
    sbe_config cfg{};
    sbe_model model{};
    std::vector<sbe_pulse> pulses;
    std::string target;
    std::string warnings;
    sbe_wannier_tb *tb = nullptr;  // owns the Wannier arrays the model points into
    int32_t shot_number = 0;       // calc.shotNumber: snapshots per period of pulse 0
    ~sbe_spec() { sbe_wannier_tb_free(tb); }
};
