// Compiles BK303's maxflow.cpp where it lies in the reference tree (include
// path set by oracle/Makefile) and adds the <int,int,long long> instantiation.
#include "maxflow.cpp"
template class Graph<int, int, long long>;
