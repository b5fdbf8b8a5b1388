/* The amle executable: the reference's main (amle.c:368-370). */
int main_amle(int argc, char *argv[]);
int main(int c, char **v) { return main_amle(c, v); }
