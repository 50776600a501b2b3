// Prints the structure-information, coordinate and energy blocks of the
// predict.x driver (aenet_amd/csrc/predict_format.h) for the Cu dimer of
// the reference's unittests/Cu-dimer-forces (first structure), with the
// values its recorded output shows; tests/test_predict.py compares the
// text with that output line for line.  Test infrastructure only.
#include "predict_format.h"

int main() {
  using namespace predict_fmt;
  Structure s;
  s.file = "dimer-x02.0.xsf";
  s.pbc = false;
  s.L[0] = 2.0002;
  s.L[4] = 0.0002;
  s.L[8] = 0.0002;
  s.names = {"Cu"};
  s.xsf_type = {1, 1};
  s.type = {2, 2};
  s.cart = {0.0001, 0.0001, 0.0001, 2.0001, 0.0001, 0.0001};
  const double F[6] = {-0.243634, 0.0, 0.0, 0.243634, 0.0, 0.0};
  print_fileinfo(1, s, 1);
  print_coordinates(0, s, s.cart.data(), F, nullptr);
  print_energy(-3.16270172, -90488.41978890, 2, F);
  return 0;
}
