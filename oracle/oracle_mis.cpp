extern "C" int oracle_mis_placeholder(void) { return 0; }
