// PanNDE::Communicator (reference modules/PanNDE/include/Communicator.hpp:78-486).  The
// device solver holds the whole mesh on one GPU, so it accepts a communicator for
// signature compatibility (nullptr included) and never calls it; only the name is needed.
#pragma once

namespace PanNDE {
class Communicator;
}  // namespace PanNDE
