/* ref_prelude.h -- force-included when compiling the reference's sources
 * (oracle/Makefile, target "ref"). The reference's src/mmlib/exception.h
 * uses the include guard __EXCEPTION_H, which libstdc++'s
 * <bits/exception.h> also uses, so whichever comes second is dropped. This
 * prelude pulls in every standard header first (so std::exception is
 * defined once), releases the shared guard, and then includes the
 * reference's own global.h, which includes its exception.h normally. No
 * reference file is modified or replaced. <climits>/<typeinfo> are
 * included because some reference files use INT_MAX and typeid without
 * including them. */
#include <climits>
#include <typeinfo>
#include <exception>
#include <new>
#include <stdexcept>
#include <string>
#include <vector>
#include <valarray>
#include <fstream>
#include <iostream>
#include <sstream>
#include <map>
#include <set>
#include <list>
#include <deque>
#include <algorithm>
#include <memory>
#include <iomanip>
#include <functional>
#include <iterator>
#include <limits>
#include <cmath>
#include <cstdlib>
#include <cstdio>
#include <cstring>
#include <ctime>
#include <math.h>
#include <float.h>
#include <ctype.h>
#include <string.h>
#include <time.h>
#include <stdio.h>
#include <stdlib.h>
#include <stdarg.h>
#undef __EXCEPTION_H
#include "global.h"
