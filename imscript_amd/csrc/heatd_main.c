/* The heatd executable: the reference's main (heatd.c:109-150). */
int main_heatd(int argc, char *argv[]);
int main(int c, char **v) { return main_heatd(c, v); }
