/* The simpois executable: the reference's main (simpois.c:441-443). */
int main_simpois(int argc, char *argv[]);
int main(int c, char **v) { return main_simpois(c, v); }
