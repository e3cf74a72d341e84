/* placeholder, filled with the growth restatement */
