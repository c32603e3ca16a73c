/*
 * common.h -- the standard headers the reference's include/common.h gathers
 * (include/common.h:4-14 of LBMFluids/sc_lbm).  Drivers and test programs
 * written against the reference reach std::cout, std::for_each, file
 * streams and the like through lbm.h / fluid.h / geometry.h, which include
 * this header, so the drop-in headers provide the same set.
 */
#ifndef SCLBM_COMMON_H
#define SCLBM_COMMON_H

#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstring>
#include <fstream>
#include <iostream>
#include <iterator>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#endif
