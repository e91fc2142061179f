// The ED programs' common main (reference: prog/main.cpp): -DPOMEROL_ANDERSON builds the Anderson
// model program, -DPOMEROL_HUBBARD2D the Hubbard cluster program.
#if defined(POMEROL_ANDERSON)
#include "anderson_model.hpp"
using model_type = anderson_model;
#elif defined(POMEROL_HUBBARD2D)
#include "hubbard2d_model.hpp"
using model_type = hubbard2d_model;
#endif

#include <exception>
#include <iostream>

int main(int argc, char* argv[]) {
    try {
        model_type m(argc, argv);
        m.compute();
    } catch(std::exception const& e) {
        std::cerr << e.what() << '\n';
        return 1;
    }
    return 0;
}
