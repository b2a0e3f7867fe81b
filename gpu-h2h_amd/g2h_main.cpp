// G2H placeholder main, replaced by the full command line below.
int main() { return 0; }
