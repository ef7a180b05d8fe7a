/* Compiler-compatibility prelude for building the reference solver's own
 * sources with g++ (the sources target MSVC 2017).  It is force-included
 * (-include) ahead of every reference translation unit by oracle/Makefile.
 *
 * The reference constructs std::exception from a string literal, e.g.
 * psstFiniteVolumeGrid.cpp:42 `throw std::exception("...")`, which only
 * MSVC's std::exception accepts.  After the standard headers have been
 * seen, `exception` is mapped to std::runtime_error, which has that
 * constructor.  Nothing else is changed: no header, library or generated
 * file is replaced.
 */
#include <exception>
#include <stdexcept>
#include <string>
#include <vector>
#include <map>
#include <memory>
#include <iostream>
#include <fstream>
#include <sstream>
#include <iomanip>
#include <cmath>
#include <chrono>
#include <typeinfo>
#include <algorithm>
#define exception runtime_error
