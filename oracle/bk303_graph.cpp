// Compiles BK303's graph.cpp where it lies in the reference tree (include path
// set by oracle/Makefile) and adds the <int,int,long long> instantiation, so
// that flows above 2^31 stay exact.  No reference source is kept in this repo.
#include "graph.cpp"
template class Graph<int, int, long long>;
