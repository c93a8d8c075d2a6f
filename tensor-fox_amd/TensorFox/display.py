"""Per-iteration tables printed for display > 1 (GaussNewton.py:140-158, 192-210;
Alternating_Least_Squares.py:94-108, 145-158): the same columns and format widths."""


def print_header(display, als=False):
    if display == 4:
        cols = ['{:^9}'.format('Iteration'), '| {:^11}'.format('Rel error'),
                '| {:^11}'.format('Step size'), '| {:^11}'.format('Improvement'),
                '| {:^11}'.format('norm(grad)')]
        if not als:
            cols += ['| {:^11}'.format('Predicted error'), '| {:^10}'.format('# Inner iterations')]
    else:
        cols = ['{:^9}'.format('Iteration'), '| {:^9}'.format('Rel error'),
                '| {:^11}'.format('Step size'), '| {:^10}'.format('Improvement'),
                '| {:^10}'.format('norm(grad)')]
        if not als:
            cols += ['| {:^10}'.format('Predicted error'), '| {:^10}'.format('# Inner iterations')]
    print('   ', *cols)


def print_row(display, it, error, step, improv, grad, predicted=None, inner=None):
    if display == 4:
        cols = ['{:^8}'.format(it + 1), '| {:^10.5e}'.format(error), '| {:^10.5e}'.format(step),
                '| {:^10.5e}'.format(improv), '| {:^11.5e}'.format(grad)]
        if predicted is not None:
            cols += ['| {:^15.5e}'.format(predicted), '| {:^16}'.format(inner)]
        print('    ', *cols)
    else:
        cols = ['{:^9}'.format(it + 1), '| {:^9.2e}'.format(error), '| {:^11.2e}'.format(step),
                '| {:^11.2e}'.format(improv), '| {:^10.2e}'.format(grad)]
        if predicted is not None:
            cols += ['| {:^15.2e}'.format(predicted), '| {:^16}'.format(inner)]
        print('   ', *cols)
