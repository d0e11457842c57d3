// bsl_reference_types.h -- the two storage types of the reference's plugin
// boundary, with the reference's layout (interfaceContainers.h:6-14), for
// translation units built outside the reference tree (the standalone driver).
// Inside the reference build, its own interfaceContainers.h is used instead.
#pragma once

#include <string>

struct lanczosStor {
    int total_basis_size;
    std::string sim_descr;
    std::string sim_descr_short;
};

struct generalStor {
    double dummy;
};
